// kernels_mixed.hip -- mixed-precision substep kernels (binned tiles and grid update).
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

template <int MODE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_tiles(StepArgs a, int rbuf, int wbuf, int record_mode)
{
    constexpr bool DO_G2P = (MODE & kTileG2P) != 0;
    constexpr bool DO_P2G = (MODE & kTileP2G) != 0;
    extern __shared__ double lds_raw[];
    __shared__ long long run_start[kMaxTX];
    __shared__ int run_pre[kMaxTX + 1];

    const int WN = a.win_w * a.win_h;
    // accumulators (FP64) first, then the FP32 velocity window
    lds_d *am = (lds_d *)lds_raw, *apx = am + WN, *apy = apx + WN;
    lds_f *vw = (lds_f *)(am + (DO_P2G ? 3 * WN : 0));   // 2*WN: (vx, vy) per window node
    const int tid = threadIdx.x;
    const int ntiles = *a.n_tiles;
    const PtRef R{(unsigned *)a.pts, a.pitch};

    for (int w = xcd_block(); w < ntiles; w += gridDim.x) {
        const TileItem ti = a.tiles[w];
        const SceneDev &S = a.scenes[ti.scene];
        const SceneConsts &c = S.c;
        const ConstsF cf = to_float(c);
        const int gridY = c.gridY;
        const int wc0 = ti.col0 - 1 - a.margin, wr0 = ti.row0 - 1 - a.margin;
        const glb_f *gr = (const glb_f *)(a.grid_mem + S.grid_off[rbuf]);
        glb_f *gw = (glb_f *)(a.grid_mem + S.grid_off[wbuf]);
        const long long gp = S.gpitch;
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
                const int col = wc0 + i / a.win_h, row = wr0 + i % a.win_h;
                float vx = 0, vy = 0;
                if (col >= 0 && col < c.ncols && row >= 0 && row < gridY) {
                    const long long gi = (long long)col * gridY + row;
                    vx = gr[kLinePxHi * gp + gi];
                    vy = gr[kLinePyHi * gp + gi];
                }
                vw[2 * i] = vx;
                vw[2 * i + 1] = vy;
            }
            if (DO_P2G) { am[i] = 0; apx[i] = 0; apy[i] = 0; }
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
#pragma unroll 1
        for (int it = 0; it < n_it; it++) {
            int k;
            if (a.lane_spread == 2)
                k = (it >> 1) * (2 * kBlock) + wave * 128 + ((lane & 31) << 2) + (lane >> 5) + 2 * (it & 1);
            else if (a.lane_spread == 1)
                k = it * kBlock + (wave << 6) + ((lane & 15) << 2) + (lane >> 4);
            else
                k = it * kBlock + tid;
            if (k >= total) continue;
            if (k < run_pre[cidx]) cidx = 0;
            while (k >= run_pre[cidx + 1]) cidx++;
            const long long pt = run_start[cidx] + (k - run_pre[cidx]);
            unsigned ud = R.u(kLineUtility, pt);
            if (ud & (unsigned)kDisabled) continue;
            unsigned xq, yq;
            float D[4], Jp, v[2], C[4];
            load_state(R, pt, xq, yq, D, Jp);
            if (DO_G2P) {
                float pp, qq;
                unsigned e;
                if (!g2p_point(c, cf, gr, gp, vw, wc0, wr0, a.win_w, a.win_h, xq, yq, v, C, D, Jp, ud, pp, qq,
                               e, (unsigned)pt, (unsigned)step)) {
                    err = max(err, e);
                    continue;
                }
                store_state(R, pt, xq, yq, v, D, C, Jp, ud, record, pp, qq);
            } else {
                v[0] = R.f(kLineVelX, pt);
                v[1] = R.f(kLineVelX + 1, pt);
#pragma unroll
                for (int m = 0; m < 4; m++) C[m] = R.f(kLineBp00 + m, pt);
            }
            if (DO_P2G) {
                const unsigned e = p2g_point(c, cf, xq, yq, v, C, D, Jp, ud,
                    [&](int col, int row, float dm, float dpx, float dpy) {
                        const int lx = col - wc0, ly = row - wr0;
                        if (lx >= 0 && lx < a.win_w && ly >= 0 && ly < a.win_h) {
                            const int li = lx * a.win_h + ly;
                            lds_add(am + li, (double)dm);
                            lds_add(apx + li, (double)dpx);
                            lds_add(apy + li, (double)dpy);
                        } else {
                            const long long gi = (long long)col * gridY + row;
                            twosum_add(gw + kLineMHi * gp + gi, gw + kLineMLo * gp + gi, dm, 0.f);
                            twosum_add(gw + kLinePxHi * gp + gi, gw + kLinePxLo * gp + gi, dpx, 0.f);
                            twosum_add(gw + kLinePyHi * gp + gi, gw + kLinePyLo * gp + gi, dpy, 0.f);
                        }
                    });
                err = max(err, e);
            }
        }
        if (err) atomicMax(&a.err[ti.scene], err);
        if (DO_P2G) {
            __syncthreads();
            for (int i = tid; i < WN; i += kBlock) {
                const double m = am[i];
                if (m == 0) continue;
                const int col = wc0 + i / a.win_h, row = wr0 + i % a.win_h;
                if (col < 0 || col >= c.ncols || row < 0 || row >= gridY) continue;
                const long long gi = (long long)col * gridY + row;
                twosum_add_d(gw + kLineMHi * gp + gi, gw + kLineMLo * gp + gi, m);
                twosum_add_d(gw + kLinePxHi * gp + gi, gw + kLinePxLo * gp + gi, apx[i]);
                twosum_add_d(gw + kLinePyHi * gp + gi, gw + kLinePyLo * gp + gi, apy[i]);
            }
        }
        __syncthreads();
    }
}

// Grid update (kernels.cuh:164-301) in FP64 on hi + lo; velocity stored in FP32.
__global__ void __launch_bounds__(kBlock) k_update(StepArgs a, int buf, int zero_other)
{
    for (int ch = blockIdx.x; ch < a.n_chunks; ch += gridDim.x) {
        const NodeChunk nc = a.chunks[ch];
        const SceneDev &S = a.scenes[nc.scene];
        const SceneConsts &c = S.c;
        const SceneClock clk = a.clocks[nc.scene];
        glb_f *g = (glb_f *)(a.grid_mem + S.grid_off[buf]);
        glb_f *go = (glb_f *)(a.grid_mem + S.grid_off[buf ^ 1]);
        double *force = a.force_mem + S.force_off;
        const long long gp = S.gpitch;
        const double dt = c.dt;
        for (int k = threadIdx.x; k < nc.count; k += kBlock) {
            const long long idx = (long long)nc.node0 + k;
            if (zero_other && go[idx] != 0) {
#pragma unroll
                for (int l = 0; l < 6; l++) go[l * gp + idx] = 0;
            }
            const float mh = g[idx];
            if (mh == 0) continue;
            const double mass = (double)mh + (double)g[kLineMLo * gp + idx];
            const double px = (double)g[kLinePxHi * gp + idx] + (double)g[kLinePxLo * gp + idx];
            const double py = (double)g[kLinePyHi * gp + idx] + (double)g[kLinePyLo * gp + idx];
            const int gix = (int)(idx / c.gridY) - c.halo + c.xoff, giy = (int)(idx % c.gridY);
            const double gnx = (double)gix * c.h, gny = (double)giy * c.h;
            double vx = px / mass, vy = py / mass;
            vy -= c.dt_gravity;
            const double sq = vx * vx + vy * vy;
            if (sq > c.vmax_sq) {
                const double r = sqrt(sq);
                vx = vx / r * c.vmax;
                vy = vy / r * c.vmax;
            }
            double n0 = gnx - clk.ind_x, n1 = gny - clk.ind_y;
            if (n0 * n0 + n1 * n1 < c.ind_rsq) {
                const double vr0 = vx - c.ind_vel, vr1 = vy;
                const double z = n0 * n0 + n1 * n1;
                if (z > 0) { const double r = sqrt(z); n0 = n0 / r; n1 = n1 / r; }
                const double vn = vr0 * n0 + vr1 * n1;
                if (vn < 0) {
                    const double pv0 = vx, pv1 = vy;
                    vx = c.ind_vel + (vr0 - n0 * vn);
                    vy = vr1 - n1 * vn;
                    const double f0 = (pv0 - vx) * mass / dt, f1 = (pv1 - vy) * mass / dt;
                    float angle = atan2f((float)n0, (float)n1);
                    angle += fp64::kPi;
                    angle *= c.n_ind_sub / (2 * fp64::kPi);
                    const int index = min(max((int)angle, 0), c.n_ind_sub - 1);
                    atomicAdd(&force[2 * index], f0);
                    atomicAdd(&force[2 * index + 1], f1);
                }
            }
            if (giy <= 2) { vx = 0; vy = 0; }
            if (giy >= c.gridY - 3 && vy > 0) vy = 0;
            if (gix <= 2 && vx < 0) vx = 0;
            else if (gix >= c.gridXTotal - 3 && vx > 0) vx = 0;
            g[kLinePxHi * gp + idx] = (float)vx;
            g[kLinePyHi * gp + idx] = (float)vy;
        }
    }
}

size_t tile_lds_bytes(const StepArgs &a, int mode)
{
    const size_t WN = (size_t)a.win_w * a.win_h;
    size_t n = 0;
    if (mode & kTileP2G) n += 3 * WN * sizeof(double);
    if (mode & kTileG2P) n += 2 * WN * sizeof(float);
    return n;
}

hipError_t launch_tiles(const StepArgs &a, int mode, int rbuf, int wbuf, int record_mode, int n_blocks,
                        hipStream_t s)
{
    if (a.tx > kMaxTX || n_blocks <= 0) return hipErrorInvalidValue;
    const size_t lds = tile_lds_bytes(a, mode);
    switch (mode) {
    case kTileG2P:
        hipLaunchKernelGGL(k_tiles<kTileG2P>, dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, record_mode);
        break;
    case kTileP2G:
        hipLaunchKernelGGL(k_tiles<kTileP2G>, dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, record_mode);
        break;
    case kTileFused:
        hipLaunchKernelGGL(k_tiles<kTileFused>, dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, record_mode);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_update(const StepArgs &a, int buf, int zero_other, int n_blocks, hipStream_t s)
{
    if (n_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_update, dim3(n_blocks), dim3(kBlock), 0, s, a, buf, zero_other);
    return hipGetLastError();
}

}  // namespace mixed
}  // namespace icempm
