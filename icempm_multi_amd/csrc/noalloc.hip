// noalloc.hip -- icempm_substeps_noalloc: substeps of one scene that allocate
// no device memory (no counterpart in the reference).
//
// For callers that cannot call hipMalloc: code under hipGraph capture, tight
// serving loops.  The caller holds the points in device memory (the
// HostSideSOA line layout, FP64) and passes them in; every other buffer the
// kernels use (the grid, its segment flags, the grid-update run table, the
// scene constants, the two clock slots, the error byte and the indenter force
// bins) lives in one file-scope device array, `scratch`, of a fixed size.  Its
// largest part is the grid, so scenes are limited to kNoAllocMaxNodes nodes.
//
// The substeps are those of icempm_partition_step with ICEMPM_STEP_DIRECT_P2G
// (unbinned P2G and G2P, one thread per point, the reference's launch shape):
// per substep the grid is cleared, P2G scatters with global atomics, the grid
// update advances the clock, G2P gathers.  The same kernels run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "engine.h"
#include "icempm_host.h"
#include "kernels.h"

namespace icempm {
namespace {

constexpr long long kNoAllocMaxNodes = ICEMPM_NOALLOC_MAX_NODES;   // (GridX + 2 halo) * GridY
constexpr long long kNoAllocMaxSegs = kNoAllocMaxNodes >> kSegShift;
constexpr long long kNoAllocMaxRuns = kNoAllocMaxSegs;          // runs of >= 1 segment
constexpr int kMaxDevices = 64;

constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }
// parts of the array, each 256-byte aligned
constexpr size_t kOffGrid = 0;
constexpr size_t kOffFlags = kOffGrid + align256(3 * kNoAllocMaxNodes * sizeof(double));
constexpr size_t kOffRuns = kOffFlags + align256(2 * kNoAllocMaxSegs * sizeof(int));
constexpr size_t kOffScene = kOffRuns + align256(kNoAllocMaxRuns * sizeof(NodeChunk));
constexpr size_t kOffClocks = kOffScene + align256(sizeof(SceneDev));
constexpr size_t kOffErr = kOffClocks + align256(2 * sizeof(SceneClock));
constexpr size_t kOffForce = kOffErr + align256(sizeof(unsigned));
constexpr size_t kScratchBytes = kOffForce + align256(2 * (size_t)kMaxIndenterSub * sizeof(double));

}  // namespace

// The scratch memory of icempm_substeps_noalloc, one copy per GPU, held until
// the process ends (about 24.2 MiB).  In host code `scratch` names a host-side
// stand-in: only the address hipGetSymbolAddress returns for a device is used.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mutex;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, looked up once per device with
// that device current (the calling thread's device is restored).
int scratch_address(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int prev = 0;
        hipError_t e = hipGetDevice(&prev);
        if (e != hipSuccess) { set_last_error(std::string("hipGetDevice: ") + hipGetErrorString(e)); return ICEMPM_ERR_HIP; }
        e = hipSetDevice(device);
        if (e != hipSuccess) { set_last_error(std::string("hipSetDevice: ") + hipGetErrorString(e)); return ICEMPM_ERR_HIP; }
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        hipSetDevice(prev);
        if (e != hipSuccess || !p) {
            set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return ICEMPM_ERR_HIP;
        }
        g_scratch_ptr[device] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[device];
    return ICEMPM_OK;
}

// Writes the scene table entry, both clock slots, the error byte, the force
// bins and the grid-update run table into the scratch array (kernel arguments
// by value: no host-to-device copy).
__global__ void k_noalloc_setup(SceneDev d, SceneClock ck, SceneDev *scene, SceneClock *clocks, unsigned *err,
                                double *force, int n_force, NodeChunk *runs, int n_runs, int run_len, int n_segs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *scene = d;
        clocks[0] = ck;
        clocks[1] = ck;
        *err = 0;
    }
    if (i < n_force) force[i] = 0;
    if (i < n_runs) {
        const int s0 = i * run_len;
        NodeChunk r;
        r.scene = 0;
        r.seg0 = s0;
        r.nseg = min(run_len, n_segs - s0);
        r.pad = 0;
        runs[i] = r;
    }
}

// Copies the error byte and the force bins to the caller's outputs.
__global__ void k_noalloc_outputs(const unsigned *err, const double *force, int n_force, unsigned *err_out,
                                  double *force_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && err_out) *err_out = *err;
    if (force_out && i < n_force) force_out[i] = force[i];
}

int bad(const char *msg)
{
    set_last_error(msg);
    return ICEMPM_ERR_ARGUMENT;
}

#define NA_HIPCHK(expr)                                                                      \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess) {                                                              \
            set_last_error(std::string(#expr) + ": " + hipGetErrorString(_e));               \
            return ICEMPM_ERR_HIP;                                                           \
        }                                                                                    \
    } while (0)

}  // namespace
}  // namespace icempm

extern "C" {

int icempm_substeps_noalloc(int device, const icempm_simparams *prms, size_t prms_bytes, double *d_points,
                            unsigned pitch, int n_points, int n_substeps, void *stream, unsigned *d_err_out,
                            double *d_force_out)
{
    using namespace icempm;
    // every argument is checked before the GPU is touched
    if (!prms) return bad("substeps_noalloc: null parameters");
    if (prms_bytes != sizeof(icempm_simparams)) return bad("substeps_noalloc: parameter struct size mismatch");
    if (n_points < 0 || n_substeps < 0) return bad("substeps_noalloc: negative count");
    if ((long long)n_points > (long long)pitch) return bad("substeps_noalloc: more points than the line pitch");
    if (n_points > 0 && !d_points) return bad("substeps_noalloc: null point buffer");
    if ((long long)pitch >= kMaxPointSpace) {
        set_last_error("substeps_noalloc: the line pitch must stay below 2^29 points");
        return ICEMPM_ERR_CAPACITY;
    }
    const icempm_simparams &P = *prms;
    if (P.GridXTotal <= 0 || P.GridY < 4 || P.GridHaloSize < 0 || !(P.cellsize > 0) || !(P.InitialTimeStep > 0) ||
        P.n_indenter_subdivisions <= 0 || P.n_indenter_subdivisions > kMaxIndenterSub)
        return bad("substeps_noalloc: parameters not derived or out of range");
    const long long nodes = ((long long)P.GridXTotal + 2LL * P.GridHaloSize) * P.GridY;
    if (nodes > kNoAllocMaxNodes) {
        set_last_error("substeps_noalloc: the grid has " + std::to_string(nodes) + " nodes (with halo columns); "
                       "this entry point holds at most " + std::to_string(kNoAllocMaxNodes));
        return ICEMPM_ERR_CAPACITY;
    }
    if (device < 0 || device >= kMaxDevices) return bad("substeps_noalloc: device index out of range");
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0) {
        set_last_error("No available HIP devices");
        return ICEMPM_ERR_HIP;
    }
    if (device >= n_dev) return bad("substeps_noalloc: device index out of range");
    if (n_substeps == 0) return ICEMPM_OK;

    unsigned char *base = nullptr;
    const int rc = scratch_address(device, &base);
    if (rc != ICEMPM_OK) return rc;
    int prev = 0;
    NA_HIPCHK(hipGetDevice(&prev));
    NA_HIPCHK(hipSetDevice(device));
    struct Restore {
        int dev;
        ~Restore() { hipSetDevice(dev); }
    } restore{prev};
    hipStream_t s = (hipStream_t)stream;

    // the scene: one, at point 0, with both grid buffers on the one grid
    // (the unbinned pipeline reads and writes buffer 0 only)
    const long long gpitch = (nodes + kSegNodes - 1) / kSegNodes * kSegNodes;
    const int n_segs = (int)(gpitch >> kSegShift);
    const int run_len = update_run_length(gpitch);
    const int n_runs = (n_segs + run_len - 1) / run_len;
    SceneDev d;
    std::memset(&d, 0, sizeof d);
    scene_consts(P, d.c);
    scene_fail_table(d.c, d.fail_tab);
    d.pt_off = 0;
    d.n_pts = n_points;
    d.cap_pts = (int)pitch;
    d.grid_off[0] = d.grid_off[1] = 0;
    d.gpitch = gpitch;
    d.ncells = (int)nodes;
    d.force_off = 0;
    d.seg_off = 0;
    SceneClock ck{};
    ck.t = P.SimulationTime;
    ck.ind_x = P.indenter_x;
    ck.ind_y = P.indenter_y;
    ck.step = P.SimulationStep;
    ck.next_step = P.SimulationStep;

    SceneDev *d_scene = (SceneDev *)(base + kOffScene);
    SceneClock *d_clocks = (SceneClock *)(base + kOffClocks);
    unsigned *d_err = (unsigned *)(base + kOffErr);
    double *d_force = (double *)(base + kOffForce);
    NodeChunk *d_runs = (NodeChunk *)(base + kOffRuns);
    int *d_flags = (int *)(base + kOffFlags);
    const int n_force = 2 * P.n_indenter_subdivisions;
    const int n_setup = std::max(n_force, n_runs);
    hipLaunchKernelGGL(k_noalloc_setup, dim3((n_setup + 255) / 256), dim3(256), 0, s, d, ck, d_scene, d_clocks,
                       d_err, d_force, n_force, d_runs, n_runs, run_len, n_segs);
    NA_HIPCHK(hipGetLastError());

    StepArgs a{};
    a.scenes = d_scene;
    a.grid_mem = (char *)(base + kOffGrid);
    a.force_mem = d_force;
    a.pts = d_points;
    a.pitch = pitch;
    a.chunks = d_runs;
    a.n_chunks = n_runs;
    a.nscenes = 1;
    a.err = d_err;
    a.dirty = d_flags;
    a.dirty_stride = n_segs;
    int n_cus = 256;
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && v > 0) n_cus = v;
    }
    const int upd_blocks = std::max(1, std::min((n_runs + 3) / 4, n_cus * 8));
    int cs = 0;
    for (int k = 0; k < n_substeps; k++) {
        StepArgs au = a, an = a;
        au.clocks = d_clocks + cs;         // the update reads this slot and writes the next clock to the other
        au.clocks_next = d_clocks + (cs ^ 1);
        an.clocks = d_clocks + (cs ^ 1);   // G2P reads the advanced clock
        an.clocks_next = d_clocks + cs;
        NA_HIPCHK(launch_fill_zero(base + kOffGrid, (size_t)3 * gpitch * sizeof(double), s));
        NA_HIPCHK(launch_fill_zero(d_flags, sizeof(int) * n_segs, s));
        NA_HIPCHK(fp64::launch_p2g_direct(au, 0, n_points, 0, s));
        NA_HIPCHK(fp64::launch_update(au, 0, 0, 1, upd_blocks, s));
        NA_HIPCHK(fp64::launch_g2p_direct(an, 0, n_points, 0, -1, s));
        cs ^= 1;
    }
    if (d_err_out || d_force_out) {
        hipLaunchKernelGGL(k_noalloc_outputs, dim3((n_force + 255) / 256), dim3(256), 0, s, d_err, d_force, n_force,
                           d_err_out, d_force_out);
        NA_HIPCHK(hipGetLastError());
    }
    return ICEMPM_OK;
}

}  // extern "C"
