// kernels_mixed_update.hip -- grid update of the mixed-precision form.
//
// Replaces partition_kernel_update_nodes (simulation/kernels.cuh:164-301) for
// the compensated (hi, lo) FP32 grid of kernels_mixed.hip.  Kept in a
// translation unit of its own so that it is compiled with the default
// instruction scheduler while the tile kernels use another one (see the
// Makefile).
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "point_mixed.h"

namespace icempm {
namespace mixed {

// Grid update (kernels.cuh:164-301) in FP64 on hi + lo; velocity stored in FP32.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) k_update(StepArgs a, int buf, int zero_other,
                                                                                            int advance_clock)
{
    // runs of 64 node segments, one per wave, as in the FP64 form; 6 FP32 lines per buffer
    constexpr int kU = 4;   // 6 FP32 lines per node: 4 segments in flight fit 4 waves per SIMD
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6), n_waves = gridDim.x * (kBlock / 64);
    for (int r = wave; r < a.n_chunks; r += n_waves) {
        const NodeChunk run = uniform_copy(a.chunks[r]);
        const SceneDev &S = a.scenes[run.scene];
        const SceneConsts c = uniform_copy(S.c);
        SceneClock clk = uniform_copy(a.clocks[run.scene]);
        if (advance_clock) clk = next_clock(c, clk);
        glb_f *g = (glb_f *)(a.grid_mem + S.grid_off[buf]);
        glb_f *go = (glb_f *)(a.grid_mem + S.grid_off[buf ^ 1]);
        double *force = a.force_mem + S.force_off;
        const long long gp = S.gpitch;
        int *fl = a.dirty + buf * a.dirty_stride + S.seg_off + run.seg0;
        int *flo = a.dirty + (buf ^ 1) * a.dirty_stride + S.seg_off + run.seg0;
        const bool in = lane < run.nseg;
        const int f = in ? fl[lane] : 0;
        const int fo = (in && zero_other) ? flo[lane] : 0;
        unsigned long long mb = __ballot(f != 0), mo = __ballot(fo != 0);
        const long long node0 = (long long)run.seg0 << kSegShift;
        while (mo) {
            const int sg = __builtin_ctzll(mo);
            mo &= mo - 1;
            const long long idx = node0 + ((long long)sg << kSegShift) + lane;
#pragma unroll
            for (int l = 0; l < 6; l++) go[l * gp + idx] = 0;
        }
        if (fo) flo[lane] = 0;
        while (mb) {
            int sg[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                sg[u] = mb ? __builtin_ctzll(mb) : -1;
                mb &= mb - 1;
            }
            float mh[kU];
#pragma unroll
            for (int u = 0; u < kU; u++)
                mh[u] = sg[u] >= 0 ? g[node0 + ((long long)sg[u] << kSegShift) + lane] : 0.f;
            float ml[kU], pxh[kU], pxl[kU], pyh[kU], pyl[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const long long idx = node0 + ((long long)sg[u] << kSegShift) + lane;
                const bool live = mh[u] != 0;
                ml[u] = live ? g[kLineMLo * gp + idx] : 0.f;
                pxh[u] = live ? g[kLinePxHi * gp + idx] : 0.f;
                pxl[u] = live ? g[kLinePxLo * gp + idx] : 0.f;
                pyh[u] = live ? g[kLinePyHi * gp + idx] : 0.f;
                pyl[u] = live ? g[kLinePyLo * gp + idx] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < kU; u++) {
                if (mh[u] == 0) continue;
                const long long idx = node0 + ((long long)sg[u] << kSegShift) + lane;
                const double mass = (double)mh[u] + (double)ml[u];
                const double px = (double)pxh[u] + (double)pxl[u];
                const double py = (double)pyh[u] + (double)pyl[u];
                const int gix = (int)(idx / c.gridY) - c.halo + c.xoff, giy = (int)(idx % c.gridY);
                double vx, vy;
                fp64::node_velocity(c, clk, force, gix, giy, mass, px, py, vx, vy);
                g[kLinePxHi * gp + idx] = (float)vx;
                g[kLinePyHi * gp + idx] = (float)vy;
            }
        }
    }
    // the advanced clocks go to the other slot, which no block reads here
    if (advance_clock && blockIdx.x == 0)
        for (int sc = threadIdx.x; sc < a.nscenes; sc += kBlock)
            a.clocks_next[sc] = next_clock(a.scenes[sc].c, a.clocks[sc]);
}

hipError_t launch_update(const StepArgs &a, int buf, int zero_other, int advance_clock, int n_blocks, hipStream_t s)
{
    if (n_blocks <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_update, dim3(n_blocks), dim3(kBlock), 0, s, a, buf, zero_other, advance_clock);
    return hipGetLastError();
}

}  // namespace mixed
}  // namespace icempm
