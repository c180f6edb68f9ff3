// kernels_mixed_direct.hip -- unbinned mixed-precision P2G and G2P (one thread
// per point; P2G adds every contribution with FP64 global atomics).  Separate
// translation unit for the same compiler reason as kernels_fp64_direct.hip.
#include <hip/hip_runtime.h>

#include "point_mixed.h"

namespace icempm {
namespace mixed {

__global__ void __launch_bounds__(kBlock) k_p2g_direct(StepArgs a, int scene, int wbuf)
{
    const SceneDev &S = a.scenes[scene];
    const SceneConsts c = uniform_copy(S.c);
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= S.n_pts) return;
    const long long pt = S.pt_off + i;
    const PtRef R{(unsigned *)a.pts, a.pitch};
    const unsigned ud = R.u(kLineUtility, pt);
    if (ud & (unsigned)kDisabled) return;
    const ConstsF cf = to_float(c);
    unsigned xq, yq;
    float D[4], Jm1, v[2], C[4];
    load_state(R, pt, xq, yq, D, Jm1);
    v[0] = R.f(kLineVelX, pt);
    v[1] = R.f(kLineVelX + 1, pt);
#pragma unroll
    for (int m = 0; m < 4; m++) C[m] = R.f(kLineBp00 + m, pt);
    glb_d *gw = (glb_d *)(a.grid_mem + S.grid_off[wbuf]);
    const long long gp = S.gpitch;
    const int gridY = c.gridY;
    const unsigned e = p2g_point(c, cf, xq, yq, v, C, D, Jm1, ud, [&](int col, int row, double dm, double dpx, double dpy) {
        const long long gi = (long long)col * gridY + row;
        grid_add(gw, gp, gi, dm, dpx, dpy);
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
    const PtRef R{(unsigned *)a.pts, a.pitch};
    unsigned ud = R.u(kLineUtility, pt);
    if (ud & (unsigned)kDisabled) return;
    const ConstsF cf = to_float(c);
    unsigned xq, yq;
    float D[4], Jm1, v[2], C[4], pp, qq;
    load_state(R, pt, xq, yq, D, Jm1);
    unsigned e;
    StressTerms terms;
    if (!g2p_point(c, cf, (const glb_d *)(a.grid_mem + S.grid_off[rbuf]), S.gpitch, nullptr, nullptr, 0, 0, 0, 0, 0,
                   false, xq, yq, v, C, D, Jm1, ud, pp, qq, e, (unsigned)pt, (unsigned)a.clocks[scene].step, terms)) {
        atomicMax(&a.err[scene], e);
        return;
    }
    const bool rec = record >= 0 ? record != 0 : (a.clocks[scene].step % c.update_every) == 0;
    store_state(R, pt, xq, yq, v, D, C, Jm1, ud, rec, pp, qq);
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

}  // namespace mixed
}  // namespace icempm
