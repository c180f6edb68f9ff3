// kernels_fp64_update.hip -- FP64 grid update for gfx950 (CDNA4).
//
// Replaces partition_kernel_update_nodes (simulation/kernels.cuh:164-301).
// Grid layout as in kernels_fp64.hip.  Kept in a translation unit of its own
// so that it is compiled with the default instruction scheduler while the
// tile kernels use another one (see the Makefile).
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "point_fp64.h"

namespace icempm {
namespace fp64 {

// ---------------------------------------------------------------------------
// Grid update (kernels.cuh:164-301, live lines), one lane per node of the
// flagged segments.  With zero_other set it also clears the other
// accumulation grid for the next P2G; with advance_clock set it computes the
// substep's clock from the current slot and block 0 writes it to the next.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) k_update(StepArgs a, int buf, int zero_other,
                                                                                            int advance_clock)
{
    // A wave takes a run of up to 64 node segments: lane l loads the flags of
    // segment l of the run for both buffers, a ballot gives the flagged ones.
    // Flagged segments of the other buffer are cleared (all 64 nodes, one per
    // lane) and their flags reset; flagged segments of this buffer are
    // updated, 8 at a time with their loads in flight together.
    constexpr int kU = 8;
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6), n_waves = gridDim.x * (kBlock / 64);
    for (int r = wave; r < a.n_chunks; r += n_waves) {
        const NodeChunk run = uniform_copy(a.chunks[r]);
        const SceneDev &S = a.scenes[run.scene];
        const SceneConsts c = uniform_copy(S.c);
        SceneClock clk = uniform_copy(a.clocks[run.scene]);
        if (advance_clock) clk = next_clock(c, clk);
        double *g = (double *)(a.grid_mem + S.grid_off[buf]);
        double *go = (double *)(a.grid_mem + S.grid_off[buf ^ 1]);
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
            go[idx] = 0;
            go[gp + idx] = 0;
            go[2 * gp + idx] = 0;
        }
        if (fo) flo[lane] = 0;
        while (mb) {
            int sg[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                sg[u] = mb ? __builtin_ctzll(mb) : -1;
                mb &= mb - 1;
            }
            double m[kU], px[kU], py[kU];
#pragma unroll
            for (int u = 0; u < kU; u++)
                m[u] = sg[u] >= 0 ? g[node0 + ((long long)sg[u] << kSegShift) + lane] : 0.0;
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const long long idx = node0 + ((long long)sg[u] << kSegShift) + lane;
                px[u] = m[u] != 0 ? g[gp + idx] : 0.0;
                py[u] = m[u] != 0 ? g[2 * gp + idx] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kU; u++) {
                if (m[u] == 0) continue;
                const long long idx = node0 + ((long long)sg[u] << kSegShift) + lane;
                const int gix = (int)(idx / c.gridY) - c.halo + c.xoff, giy = (int)(idx % c.gridY);
                double vx, vy;
                node_velocity(c, clk, force, gix, giy, m[u], px[u], py[u], vx, vy);
                g[gp + idx] = vx;
                g[2 * gp + idx] = vy;
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

}  // namespace fp64
}  // namespace icempm
