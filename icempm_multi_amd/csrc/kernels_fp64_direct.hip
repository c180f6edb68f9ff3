// kernels_fp64_direct.hip -- unbinned FP64 P2G and G2P: one thread per point,
// 27 global FP64 atomics per point in P2G (the reference's launch shape,
// gpu_partition.cu:230-293).  Used before the first sort and as the baseline.
// Kept in its own translation unit: compiled together with the binned kernels
// it trips an inliner crash of the ROCm 7.2 device compiler.
#include <hip/hip_runtime.h>

#include "point_fp64.h"

namespace icempm {
namespace fp64 {

// ---------------------------------------------------------------------------
// Unbinned kernels: one thread per point, 27 global FP64 atomics per point.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_p2g_direct(StepArgs a, int scene, int wbuf)
{
    const SceneDev &S = a.scenes[scene];
    const SceneConsts c = uniform_copy(S.c);
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= S.n_pts) return;
    const long long pt = S.pt_off + i;
    const PtRef R{(double *)a.pts, a.pitch};
    const long long ud = as_ll(R.at(kLineUtility, pt));
    if (ud & kDisabled) return;
    double x, y, F[4], Jp, v[2], B[4];
    load_state(R, pt, x, y, F, Jp);
    v[0] = R.at(kLineVelX, pt);
    v[1] = R.at(kLineVelX + 1, pt);
#pragma unroll
    for (int m = 0; m < 4; m++) B[m] = R.at(kLineBp00 + m, pt);
    glb_d *gw = (glb_d *)(a.grid_mem + S.grid_off[wbuf]);
    const long long gp = S.gpitch;
    const int gridY = c.gridY;
    const unsigned e = p2g_point(c, x, y, v, B, F, Jp, ud, [&](int col, int row, double dm, double dpx, double dpy) {
        const long long gi = (long long)col * gridY + row;
        glb_add(gw + gi, dm);
        glb_add(gw + gp + gi, dpx);
        glb_add(gw + 2 * gp + gi, dpy);
        mark_dirty(a, S, wbuf, gi);
    });
    if (e) atomicMax(&a.err[scene], e);
}

__global__ void __launch_bounds__(kBlock) k_g2p_direct(StepArgs a, int scene, int rbuf, int record)
{
    const SceneDev &S = a.scenes[scene];
    const SceneConsts c = uniform_copy(S.c);
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= S.n_pts) return;
    const long long pt = S.pt_off + i;
    const PtRef R{(double *)a.pts, a.pitch};
    long long ud = as_ll(R.at(kLineUtility, pt));
    if (ud & kDisabled) return;
    double x, y, F[4], Jp, v[2], B[4], pp, qq;
    load_state(R, pt, x, y, F, Jp);
    unsigned e;
    StressTerms terms;
    if (!g2p_point(c, (const glb_d *)S.fail_tab, (const glb_d *)(a.grid_mem + S.grid_off[rbuf]), S.gpitch, nullptr,
                   nullptr, 0, 0, 0, 0, 0, false, x, y, v, B, F, Jp, ud, pp, qq, e, terms)) {
        atomicMax(&a.err[scene], e);
        return;
    }
    const bool rec = record >= 0 ? record != 0 : (a.clocks[scene].step % c.update_every) == 0;
    store_state(R, pt, x, y, v, F, B, Jp, ud, rec, pp, qq);
}

hipError_t launch_p2g_direct(const StepArgs &a, int scene, int n_pts, int wbuf, hipStream_t s)
{
    if (n_pts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_p2g_direct, dim3((n_pts + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a, scene, wbuf);
    return hipGetLastError();
}

hipError_t launch_g2p_direct(const StepArgs &a, int scene, int n_pts, int rbuf, int record_pq, hipStream_t s)
{
    if (n_pts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_g2p_direct, dim3((n_pts + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a, scene, rbuf,
                       record_pq);
    return hipGetLastError();
}

}  // namespace fp64
}  // namespace icempm
