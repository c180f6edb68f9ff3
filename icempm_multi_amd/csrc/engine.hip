// engine.hip -- the substep pipeline on one device (host code).
//
// Substep sequence of icy::Model::Step (model.cpp:38-54):
//   t += dt; indenter_x = x0 + t*V; reset_grid; p2g; update_nodes(t); g2p(recordPQ)
// Binned pipeline, three accumulation grids per scene (indices mod 3):
//   [P2G -> 0]; [update 0 + G2P from 0 + P2G -> 1, clear 2]; [update 1 + G2P
//   from 1 + P2G -> 2, clear 0]; ...; [update + G2P of the last substep]
// one tile launch per substep, so each substep after the first reads the
// points once and writes them once.
// Unbinned fallback (before the first sort): zero A; P2G direct; update(A); G2P direct.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"
#include "icempm_host.h"

namespace icempm {

// Tile candidates of the autotuner: {tx, ty, lane spread, dynamic queues}.
// The widths between the powers of two let the tile count of a scene come
// close to a whole number of rounds of the resident workgroups: a launch
// whose last round holds a few tiles runs that round with most of the GPU
// idle (on the 4M-point sheet, 8x64 tiles make 2.6 rounds, and the launch
// costs 12% more per point than at a tile count of 2.0 rounds).
const int kTileCandidates[][4] = {{4, 16, 1, 0},  {8, 8, 1, 0},   {4, 32, 1, 0},  {4, 64, 1, 0},  {8, 16, 1, 0},
                                  {8, 32, 1, 0},  {16, 16, 1, 0}, {4, 32, 0, 0},  {8, 32, 0, 0},  {4, 32, 2, 0},
                                  {8, 32, 2, 0},  {16, 32, 1, 0}, {8, 64, 1, 0},  {32, 16, 1, 0}, {5, 64, 1, 0},
                                  {6, 64, 1, 0},  {7, 64, 1, 0},  {10, 64, 1, 0}, {6, 32, 1, 0},  {10, 32, 1, 0},
                                  {12, 32, 1, 0}, {14, 32, 1, 0}, {20, 32, 1, 0}, {24, 32, 1, 0}, {24, 16, 1, 0}};
const int kNumTileCandidates = (int)(sizeof(kTileCandidates) / sizeof(kTileCandidates[0]));

namespace {

inline long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

int bits_for(long long n)   // smallest b with n <= 2^b
{
    int b = 0;
    while ((1LL << b) < n) b++;
    return b;
}

}  // namespace

Engine::Engine(int device, int precision) : device_(device), precision_(precision)
{
    hipSetDevice(device_);
    hipStreamCreate(&stream_);
    hipEventCreate(&ev_a_);
    hipEventCreate(&ev_b_);
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device_) == hipSuccess && n > 0) n_cus_ = n;
}

Engine::~Engine()
{
    hipSetDevice(device_);
    if (stream_) hipStreamSynchronize(stream_);
    invalidate_graphs();
    for (auto &e : ev_pool_) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    if (ev_a_) hipEventDestroy(ev_a_);
    if (ev_b_) hipEventDestroy(ev_b_);
    void *bufs[] = {pts_[0], pts_[1], grid_mem_, ind_force_mem_, d_scenes_, d_clocks_, d_err_, d_chunks_,
                    bin_mem_, tile_mem_, d_tile_queue_, d_dirty_, d_defer_};
    for (void *b : bufs)
        if (b) hipFree(b);
    if (stream_) hipStreamDestroy(stream_);
}

int Engine::fail(int code, const std::string &msg)
{
    set_last_error(msg);
    return code;
}

int Engine::hipfail(hipError_t e, const char *what)
{
    set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return ICEMPM_ERR_HIP;
}

#define HIPCHK(expr)                                           \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return hipfail(_e, #expr);       \
    } while (0)

#define RC(expr)                                               \
    do {                                                       \
        int _rc = (expr);                                      \
        if (_rc != ICEMPM_OK) return _rc;                      \
    } while (0)

long long Engine::total_points() const
{
    long long n = 0;
    for (auto &s : scenes_) n += s.n_pts;
    return n;
}

void Engine::clock(int scene, double *t, int *step) const
{
    *t = scenes_[scene].sim_time;
    *step = scenes_[scene].sim_step;
}

int Engine::add_scene(const icempm_simparams *prms, int cap_pts)
{
    if (cap_pts <= 0) return fail(ICEMPM_ERR_ARGUMENT, "add_scene: capacity must be positive");
    SceneHost s;
    s.cap_pts = cap_pts;
    if (prms) {
        s.prms = *prms;
        s.have_prms = true;
        s.sim_time = prms->SimulationTime;
        s.sim_step = prms->SimulationStep;
    }
    scenes_.push_back(s);
    committed_ = false;
    return (int)scenes_.size() - 1;
}

int Engine::set_params(int scene, const icempm_simparams &P)
{
    if (scene < 0 || scene >= nscenes()) return fail(ICEMPM_ERR_ARGUMENT, "set_params: bad scene");
    if (P.GridXTotal <= 0 || P.GridY < 4 || P.GridHaloSize < 0 || !(P.cellsize > 0) || !(P.InitialTimeStep > 0) ||
        P.n_indenter_subdivisions <= 0 || P.n_indenter_subdivisions > kMaxIndenterSub)
        return fail(ICEMPM_ERR_ARGUMENT, "update_constants: parameters not derived or out of range");
    SceneHost &s = scenes_[scene];
    const bool geometry_changed = !s.have_prms || s.prms.GridXTotal != P.GridXTotal || s.prms.GridY != P.GridY ||
                                  s.prms.GridHaloSize != P.GridHaloSize ||
                                  s.prms.n_indenter_subdivisions != P.n_indenter_subdivisions;
    const bool position_map_changed =
        precision_ == ICEMPM_PRECISION_MIXED && s.have_prms && s.prms.cellsize != P.cellsize;
    if (position_map_changed && s.n_pts > 0)
        return fail(ICEMPM_ERR_STATE, "update_constants: the mixed form cannot change the cell size of uploaded points");
    s.prms = P;
    s.have_prms = true;
    s.sim_time = P.SimulationTime;
    s.sim_step = P.SimulationStep;
    invalidate_graphs();
    if (geometry_changed && committed_) {
        // grids depend on the geometry: rebuild everything but keep the points
        committed_ = false;
        bins_valid_ = false;
    }
    if (committed_) {
        RC(push_scenes());
        RC(push_clocks());
        HIPCHK(hipMemsetAsync(d_err_, 0, sizeof(unsigned) * nscenes(), stream_));
    }
    return ICEMPM_OK;
}

int Engine::set_grid_span(int scene, int xoff, int gx)
{
    if (scene < 0 || scene >= nscenes() || !scenes_[scene].have_prms)
        return fail(ICEMPM_ERR_STATE, "set_grid_span: parameters not set");
    SceneHost &s = scenes_[scene];
    if (xoff != 0 || gx != s.prms.GridXTotal)
        return fail(ICEMPM_ERR_ARGUMENT,
                    "set_grid_span: one partition spans the whole grid on one device (GridX_offset 0, "
                    "GridX_partition = GridXTotal)");
    return ICEMPM_OK;
}

// Fixed-point position map of the mixed form: x/h = xq * 2^-fb - origin.
static int frac_bits_for(const icempm_simparams &P, int &origin)
{
    const int halo = P.GridHaloSize;
    const int ncols = P.GridXTotal + 2 * halo;
    origin = halo + 2;
    const long long span = std::max<long long>(ncols + 4, P.GridY + 4);
    // fx = frac * 2^-fb must stay exact in FP32 (frac < 1.5 * 2^fb): fb <= 23
    return std::min(23, 32 - bits_for(span));
}

void scene_consts(const icempm_simparams &P, SceneConsts &c)
{
    std::memset(&c, 0, sizeof c);
    c.gridX = P.GridXTotal;
    c.gridY = P.GridY;
    c.halo = P.GridHaloSize;
    c.ncols = c.gridX + 2 * c.halo;
    c.xoff = 0;
    c.gridXTotal = P.GridXTotal;
    c.n_ind_sub = P.n_indenter_subdivisions;
    c.update_every = std::max(1, P.UpdateEveryNthStep);
    int origin = 0;
    c.frac_bits = frac_bits_for(P, origin);
    c.origin = origin;
    c.origin_y = 2;
    c.h = P.cellsize;
    c.hinv = P.cellsize_inv;
    c.dt = P.InitialTimeStep;
    c.dt_vol_Dpinv = P.dt_vol_Dpinv;
    c.dt_gravity = P.dt_Gravity;
    c.vmax = P.vmax;
    c.vmax_sq = P.vmax_squared;
    c.pmass = P.ParticleMass;
    c.wmass = P.WaterParticleMass;
    c.mu = P.mu;
    c.kappa = P.kappa;
    c.ics = P.IceCompressiveStrength;
    c.its = P.IceTensileStrength;
    c.iss = P.IceShearStrength;
    c.its2 = P.IceTensileStrength2;
    c.beta = P.NACC_beta;
    c.grain_var = P.GrainVariability;
    c.tan_phi = P.DP_tan_phi;
    c.dp_threshold_p = P.DP_threshold_p;
    c.ind_rsq = P.IndRSq;
    c.ind_vel = P.IndVelocity;
    c.ind_x0 = P.indenter_x_initial;
    c.ind_y = P.indenter_y;
    // scene invariants of the point loop, with the operations the reference
    // applies per point (kernels.cuh:30-33, 552-555, 560-576, 598-619, 636-638)
    {
        const double pmax = c.ics, pmin = -c.its;
        c.s4 = 4. * c.hinv;
        c.kappa_half = c.kappa * 0.5;
        c.coeff1_mu = 1.4142135623730950 * c.mu;
        c.gv033 = c.grain_var * 0.033;
        c.gv01 = c.grain_var * 0.1;
        c.b2 = 1 + 2 * c.beta;
        c.bpmax = c.beta * pmax;
        c.dp = pmax - pmin;
        c.dp2 = (pmax - pmin) * (pmax - pmin);
        c.q0a = 2 * std::sqrt(-pmax * pmin);
        const double p_new = -c.dp_threshold_p;
        c.je_tension = std::sqrt(-2. * p_new / c.kappa + 1.);
        c.sq_je_tension = std::sqrt(c.je_tension);
    }
    ConstsF &f = c.f;
    f.h = (float)c.h;
    f.hinv = (float)c.hinv;
    f.dt = (float)c.dt;
    f.dt_vol_Dpinv = (float)c.dt_vol_Dpinv;
    f.pmass = (float)c.pmass;
    f.wmass = (float)c.wmass;
    f.mu = (float)c.mu;
    f.kappa = (float)c.kappa;
    f.ics = (float)c.ics;
    f.its = (float)c.its;
    f.iss = (float)c.iss;
    f.its2 = (float)c.its2;
    f.beta = (float)c.beta;
    f.grain_var = (float)c.grain_var;
    f.s4 = (float)(4. * c.hinv);
    f.dq_scale = (float)(c.dt * c.hinv * (double)(1LL << c.frac_bits));
    f.fb = c.frac_bits;
    f.origin = c.origin;
    f.origin_y = c.origin_y;
    {
        const double pmax = c.ics, pmin = -c.its, dp = pmax - pmin;
        f.gv033 = (float)(c.grain_var * 0.033);
        f.gv01 = (float)(c.grain_var * 0.1);
        f.q0c = (float)(2. * std::sqrt(-pmax * pmin) / dp);
        f.msq_c = (float)(4. * (1. + 2. * c.beta) / (dp * dp));
        f.b2 = (float)(1. + 2. * c.beta);
        f.bpmax = (float)(c.beta * pmax);
    }
}

void scene_fail_table(const SceneConsts &c, double *tab)
{
#pragma clang fp contract(off)
    const double pmax = c.ics, pmin = -c.its;
    for (int r = 0; r < kFailTab11; r++) {
        // var3 = 1 + GrainVariability * 0.1 * (-10 + (grain + 4) % 11)
        const double t = c.gv01 * (double)(-10 + (r - 10));
        const double var3 = 1.0 + t;
        const double qmax = c.iss * var3;
        tab[2 * r] = (4 * qmax * qmax * c.b2) / ((pmax - pmin) * (pmax - pmin));   // mSq
        tab[2 * r + 1] = c.q0a * qmax / (pmax - pmin);                               // q0
    }
    for (int s = 0; s < kFailTab30; s++) {
        // var2 = 1 + GrainVariability * 0.033 * (-15 + (grain + 3) % 30)
        const double t = c.gv033 * (double)(-15 + (s - 29));
        const double var2 = 1.0 + t;
        tab[2 * kFailTab11 + s] = -c.its2 * var2;                                     // pmin2
    }
}

int Engine::build_scene_dev(int i, SceneDev &d) const
{
    const SceneHost &s = scenes_[i];
    std::memset(&d, 0, sizeof d);
    scene_consts(s.prms, d.c);
    scene_fail_table(d.c, d.fail_tab);
    const SceneConsts &c = d.c;
    d.pt_off = s.pt_off;
    d.n_pts = s.n_pts;
    d.cap_pts = s.cap_pts;
    for (int b = 0; b < kGridBuffers; b++) d.grid_off[b] = (char *)s.grid[b] - (char *)grid_mem_;
    d.gpitch = s.grid_elems;
    d.cell_off = s.cells_off;
    d.ncells = c.ncols * c.gridY;
    d.n_tiles_x = (c.ncols + tx_ - 1) / tx_;
    d.n_tiles_y = (c.gridY + ty_ - 1) / ty_;
    d.force_off = s.ind_force - ind_force_mem_;
    d.seg_off = s.seg_off;
    return ICEMPM_OK;
}

int Engine::push_scenes()
{
    host_dev_.resize(nscenes());
    for (int i = 0; i < nscenes(); i++) build_scene_dev(i, host_dev_[i]);
    HIPCHK(hipMemcpyAsync(d_scenes_, host_dev_.data(), sizeof(SceneDev) * nscenes(), hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ICEMPM_OK;
}

int Engine::push_clocks()
{
    std::vector<SceneClock> ck(nscenes());
    for (int i = 0; i < nscenes(); i++) {
        const SceneHost &s = scenes_[i];
        ck[i].t = s.sim_time;
        ck[i].ind_x = s.prms.indenter_x;
        ck[i].ind_y = s.prms.indenter_y;
        ck[i].step = s.sim_step;
        ck[i].next_step = s.sim_step;
    }
    HIPCHK(hipMemcpyAsync(d_clocks_ + clock_slot_ * nscenes(), ck.data(), sizeof(SceneClock) * nscenes(),
                          hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ICEMPM_OK;
}

void Engine::invalidate_graphs()
{
    for (auto &g : graphs_) {
        hipGraphExecDestroy(g.second.exec);
        for (auto &e : g.second.ev) {
            hipEventDestroy(e.first);
            hipEventDestroy(e.second);
        }
    }
    graphs_.clear();
}

// Allocates all device memory for the current scene list.  Points already on
// the device are preserved when only the grids change.
int Engine::commit()
{
    hipSetDevice(device_);
    if (scenes_.empty()) return fail(ICEMPM_ERR_STATE, "commit: no scene");
    for (auto &s : scenes_)
        if (!s.have_prms) return fail(ICEMPM_ERR_STATE, "commit: parameters of a scene are not set");
    invalidate_graphs();
    HIPCHK(hipStreamSynchronize(stream_));
    const int eb = elem_bytes();

    // points (kept if the layout is unchanged)
    long long off = 0;
    std::vector<long long> offs(nscenes());
    for (int i = 0; i < nscenes(); i++) {
        offs[i] = off;
        off += round_up(scenes_[i].cap_pts, 256);
    }
    const long long space = off;
    // the kernels address a point line with a 32-bit byte offset (index * 8)
    if (round_up(space, 256) >= kMaxPointSpace)
        return fail(ICEMPM_ERR_CAPACITY, "point capacity of all scenes exceeds 2^29 - 256 points");
    bool points_layout_same = pts_[0] && space == n_pts_space_;
    for (int i = 0; i < nscenes() && points_layout_same; i++) points_layout_same = offs[i] == scenes_[i].pt_off;
    if (!points_layout_same) {
        for (int b = 0; b < 2; b++)
            if (pts_[b]) { hipFree(pts_[b]); pts_[b] = nullptr; }
        n_pts_space_ = space;
        pitch_ = round_up(space, 256);
        for (int b = 0; b < 2; b++) {
            HIPCHK(hipMalloc(&pts_[b], (size_t)kNumPtsLines * pitch_ * eb));
            HIPCHK(hipMemsetAsync(pts_[b], 0, (size_t)kNumPtsLines * pitch_ * eb, stream_));
        }
        for (int i = 0; i < nscenes(); i++) { scenes_[i].pt_off = offs[i]; scenes_[i].n_pts = 0; }
    }

    // deferred-point list of the kTileFromMomenta launches: a point index and
    // a scene per point of the point space, and two counters (zero)
    if (!points_layout_same || !d_defer_) {
        if (d_defer_) { hipFree(d_defer_); d_defer_ = nullptr; }
        HIPCHK(hipMalloc(&d_defer_, sizeof(int) * (2 * n_pts_space_ + 64)));
        HIPCHK(hipMemsetAsync(d_defer_ + 2 * n_pts_space_, 0, sizeof(int) * 64, stream_));
    }

    // grids: [buffer 0 of all scenes][buffer 1 ...][buffer 2 ...], 3 FP64 lines
    // (m, px, py) per scene in both precision forms, 24 bytes per node
    if (grid_mem_) { hipFree(grid_mem_); grid_mem_ = nullptr; }
    long long gelems = 0, cells = 0, forces = 0;
    std::vector<long long> goff(nscenes()), coff(nscenes()), foff(nscenes());
    const long long lines = kGridLines;
    for (int i = 0; i < nscenes(); i++) {
        const icempm_simparams &P = scenes_[i].prms;
        const long long ncols = P.GridXTotal + 2LL * P.GridHaloSize;
        const long long nodes = ncols * P.GridY;
        scenes_[i].grid_elems = round_up(nodes, 64);
        goff[i] = gelems;
        gelems += lines * scenes_[i].grid_elems;
        coff[i] = cells;
        cells += nodes + 1;
        foff[i] = forces;
        forces += 2LL * P.n_indenter_subdivisions;
    }
    // cell keys, bin counts and cell starts are 32-bit
    if (cells >= (1LL << 31) - 1) return fail(ICEMPM_ERR_CAPACITY, "grid cells of all scenes exceed 2^31");
    const size_t buf_bytes = (size_t)gelems * sizeof(double);
    HIPCHK(hipMalloc(&grid_mem_, kGridBuffers * buf_bytes));
    HIPCHK(hipMemsetAsync(grid_mem_, 0, kGridBuffers * buf_bytes, stream_));
    // segment flags (grid_elems is a multiple of the segment size); each scene's
    // range starts 16-byte aligned for the zero-fill kernel
    n_segs_ = 0;
    for (int i = 0; i < nscenes(); i++) {
        scenes_[i].seg_off = n_segs_;
        n_segs_ += round_up(scenes_[i].grid_elems >> kSegShift, 4);
    }
    if (d_dirty_) { hipFree(d_dirty_); d_dirty_ = nullptr; }
    HIPCHK(hipMalloc(&d_dirty_, sizeof(int) * kGridBuffers * std::max<long long>(n_segs_, 1)));
    HIPCHK(hipMemsetAsync(d_dirty_, 0, sizeof(int) * kGridBuffers * std::max<long long>(n_segs_, 1), stream_));
    if (ind_force_mem_) hipFree(ind_force_mem_);
    HIPCHK(hipMalloc(&ind_force_mem_, sizeof(double) * std::max<long long>(forces, 1)));
    HIPCHK(hipMemsetAsync(ind_force_mem_, 0, sizeof(double) * std::max<long long>(forces, 1), stream_));
    for (int i = 0; i < nscenes(); i++) {
        for (int b = 0; b < kGridBuffers; b++) scenes_[i].grid[b] = (char *)grid_mem_ + b * buf_bytes + goff[i] * sizeof(double);
        scenes_[i].cells_off = coff[i];
        scenes_[i].ind_force = ind_force_mem_ + foff[i];
        for (int b = 0; b < kGridBuffers; b++) scenes_[i].clean[b] = true;
    }

    // tables
    if (d_scenes_) hipFree(d_scenes_);
    if (d_clocks_) hipFree(d_clocks_);
    if (d_err_) hipFree(d_err_);
    HIPCHK(hipMalloc(&d_scenes_, sizeof(SceneDev) * nscenes()));
    HIPCHK(hipMalloc(&d_clocks_, sizeof(SceneClock) * 2 * nscenes()));   // two slots
    clock_slot_ = 0;
    HIPCHK(hipMalloc(&d_err_, sizeof(unsigned) * nscenes()));
    HIPCHK(hipMemsetAsync(d_err_, 0, sizeof(unsigned) * nscenes(), stream_));
    if (!d_tile_queue_) {
        // tile queue counters: zero here, and cleared again by the last block of every launch
        HIPCHK(hipMalloc(&d_tile_queue_, sizeof(int) * 16));
        HIPCHK(hipMemsetAsync(d_tile_queue_, 0, sizeof(int) * 16, stream_));
    }

    std::vector<NodeChunk> chunks;
    for (int i = 0; i < nscenes(); i++) {
        const int nseg = (int)(scenes_[i].grid_elems >> kSegShift);
        const int len = update_run_length(scenes_[i].grid_elems);
        for (int s0 = 0; s0 < nseg; s0 += len) chunks.push_back({i, s0, std::min(len, nseg - s0), 0});
    }
    if (d_chunks_) hipFree(d_chunks_);
    n_chunks_ = (int)chunks.size();
    HIPCHK(hipMalloc(&d_chunks_, sizeof(NodeChunk) * chunks.size()));
    HIPCHK(hipMemcpy(d_chunks_, chunks.data(), sizeof(NodeChunk) * chunks.size(), hipMemcpyHostToDevice));

    bins_.n_cells_total = cells;
    bins_.n_pts_total = n_pts_space_;
    RC(alloc_bins());
    committed_ = true;
    bins_valid_ = false;
    RC(push_scenes());
    RC(push_clocks());
    HIPCHK(hipStreamSynchronize(stream_));
    return ICEMPM_OK;
}

int Engine::alloc_bins()
{
    if (bin_mem_) { hipFree(bin_mem_); bin_mem_ = nullptr; }
    const long long np = std::max<long long>(n_pts_space_, 1), nc = std::max<long long>(bins_.n_cells_total, 1);
    const long long nscan = std::max(np, nc) / 1000 + 64;
    const long long n_int = 3 * np + 3 * nc + nscan + nscenes() + 16;
    HIPCHK(hipMalloc(&bin_mem_, sizeof(int) * n_int));
    int *p = (int *)bin_mem_;
    bins_.keys = p; p += np;
    bins_.perm = p; p += np;
    bins_.sorted_keys = p; p += np;
    bins_.cell_count = p; p += nc;
    bins_.cell_fill = p; p += nc;
    bins_.cell_start = p; p += nc;
    bins_.scan_tmp = p; p += nscan;
    bins_.scene_count = p;
    HIPCHK(hipMemsetAsync(bin_mem_, 0, sizeof(int) * n_int, stream_));
    // tile table for the current shape
    if (tile_mem_) { hipFree(tile_mem_); tile_mem_ = nullptr; }
    long long mt = 0;
    for (auto &s : scenes_) {
        const long long ncols = s.prms.GridXTotal + 2LL * s.prms.GridHaloSize;
        mt += ((ncols + tx_ - 1) / tx_) * ((s.prms.GridY + ty_ - 1) / ty_);
    }
    max_tiles_ = mt;
    const size_t tbytes = sizeof(TileItem) * mt + sizeof(int) * (4 * (mt + nscenes() + 1) + mt / 1000 + 64) + 64;
    HIPCHK(hipMalloc(&tile_mem_, tbytes));
    HIPCHK(hipMemsetAsync(tile_mem_, 0, tbytes, stream_));
    d_tiles_ = (TileItem *)tile_mem_;
    d_n_tiles_ = (int *)((char *)tile_mem_ + tbytes - 64);
    return ICEMPM_OK;
}

int Engine::ensure_committed()
{
    if (committed_) return ICEMPM_OK;
    return commit();
}

StepArgs Engine::args() const
{
    StepArgs a{};
    a.scenes = d_scenes_;
    a.grid_mem = (char *)grid_mem_;
    a.force_mem = ind_force_mem_;
    a.clocks = d_clocks_ + clock_slot_ * nscenes();
    a.clocks_next = d_clocks_ + (clock_slot_ ^ 1) * nscenes();
    a.pts = pts_[0];
    a.pitch = pitch_;
    a.cell_start = bins_.cell_start;
    a.tiles = d_tiles_;
    a.n_tiles = d_n_tiles_;
    a.chunks = d_chunks_;
    a.n_chunks = n_chunks_;
    a.nscenes = nscenes();
    a.err = d_err_;
    a.tx = tx_;
    a.ty = ty_;
    a.margin = margin_;
    a.win_w = tx_ + 2 + 2 * margin_;
    a.win_h = ty_ + 2 + 2 * margin_;
    a.lane_spread = lane_spread_;
    a.dyn_tiles = dyn_tiles_;
    a.tile_queue = d_tile_queue_;
    a.dirty = d_dirty_;
    a.defer_pts = d_defer_;
    a.defer_scene = d_defer_ ? d_defer_ + n_pts_space_ : nullptr;
    a.defer_count = d_defer_ ? d_defer_ + 2 * n_pts_space_ : nullptr;
    a.dirty_stride = n_segs_;
    return a;
}

int Engine::set_tile_shape(int tx, int ty, int margin)
{
    if (tx < 1 || tx > 32 || ty < 1 || ty > 1024 || margin < 0 || margin > 4)
        return fail(ICEMPM_ERR_ARGUMENT, "tile shape out of range");
    // LDS of the fused launch per window node: (vx, vy) and two FP64 (m, px, py)
    // accumulators, 64 B in FP64 and 56 B in the mixed form (FP32 velocities)
    const size_t per_node = precision_ == ICEMPM_PRECISION_MIXED ? 56 : 64;
    const size_t lds = (size_t)(tx + 2 + 2 * margin) * (ty + 2 + 2 * margin) * per_node;
    if (lds > 64 * 1024) return fail(ICEMPM_ERR_ARGUMENT, "tile window exceeds the LDS budget");
    if (tx == tx_ && ty == ty_ && margin == margin_) return ICEMPM_OK;
    for (int &r : resident_) r = 0;
    tx_ = tx;
    ty_ = ty;
    margin_ = margin;
    invalidate_graphs();
    if (!committed_) return ICEMPM_OK;
    // the tile table depends on the shape; cell_start does not
    if (tile_mem_) { hipFree(tile_mem_); tile_mem_ = nullptr; }
    long long mt = 0;
    for (auto &s : scenes_) {
        const long long ncols = s.prms.GridXTotal + 2LL * s.prms.GridHaloSize;
        mt += ((ncols + tx_ - 1) / tx_) * ((s.prms.GridY + ty_ - 1) / ty_);
    }
    max_tiles_ = mt;
    const size_t tbytes = sizeof(TileItem) * mt + sizeof(int) * (4 * (mt + nscenes() + 1) + mt / 1000 + 64) + 64;
    HIPCHK(hipMalloc(&tile_mem_, tbytes));
    HIPCHK(hipMemsetAsync(tile_mem_, 0, tbytes, stream_));
    d_tiles_ = (TileItem *)tile_mem_;
    d_n_tiles_ = (int *)((char *)tile_mem_ + tbytes - 64);
    RC(push_scenes());
    if (bins_valid_) RC(build_tiles());
    return ICEMPM_OK;
}

// ---------------------------------------------------------------------------
// points in and out
// ---------------------------------------------------------------------------
int Engine::upload(int scene, const double *hssoa, unsigned capacity, int offset, int n)
{
    RC(ensure_committed());
    SceneHost &s = scenes_[scene];
    if (n < 0 || n > s.cap_pts) return fail(ICEMPM_ERR_CAPACITY, "transfer_points_from_soa_to_device: capacity");
    if (offset < 0 || (long long)offset + n > (long long)capacity)
        return fail(ICEMPM_ERR_CAPACITY, "transfer_points_from_soa_to_device: HSSOA capacity");
    if (precision_ == ICEMPM_PRECISION_FP64) {
        for (int line = 0; line < kNumPtsLines && n > 0; line++)
            HIPCHK(hipMemcpyAsync((double *)pts_[0] + line * pitch_ + s.pt_off, hssoa + (size_t)line * capacity + offset,
                                  sizeof(double) * n, hipMemcpyHostToDevice, stream_));
    } else {
        const SceneDev &d = host_dev_[scene];
        const double scale = std::ldexp(1.0, d.c.frac_bits);
        const double lim = 4294967295.0;
        std::vector<uint32_t> buf((size_t)kNumPtsLines * n);
        auto L = [&](int line, int i) { return hssoa[(size_t)line * capacity + offset + i]; };
        for (int i = 0; i < n; i++) {
            long long ud;
            std::memcpy(&ud, hssoa + offset + i, 8);
            buf[(size_t)0 * n + i] = (uint32_t)ud;
            const int org[2] = {d.c.origin, d.c.origin_y};
            for (int k = 0; k < 2; k++) {
                double q = std::nearbyint((L(kLinePosX + k, i) * d.c.hinv + org[k]) * scale);
                q = std::min(std::max(q, 0.0), lim);
                buf[(size_t)(kLinePosX + k) * n + i] = (uint32_t)q;
            }
            auto putf = [&](int line, double v) {
                const float f = (float)v;
                std::memcpy(&buf[(size_t)line * n + i], &f, 4);
            };
            putf(kLineP, L(kLineP, i));
            putf(kLineQ, L(kLineQ, i));
            putf(kLineJp, L(kLineJp, i) - 1.0);   // the mixed form stores Jp_inv - 1
            putf(kLineVelX, L(kLineVelX, i));
            putf(kLineVelX + 1, L(kLineVelX + 1, i));
            putf(kLineFe00, L(kLineFe00, i) - 1.0);
            putf(kLineFe00 + 1, L(kLineFe00 + 1, i));
            putf(kLineFe00 + 2, L(kLineFe00 + 2, i));
            putf(kLineFe00 + 3, L(kLineFe00 + 3, i) - 1.0);
            for (int k = 0; k < 4; k++) putf(kLineBp00 + k, L(kLineBp00 + k, i));
        }
        for (int line = 0; line < kNumPtsLines && n > 0; line++)
            HIPCHK(hipMemcpyAsync((uint32_t *)pts_[0] + line * pitch_ + s.pt_off, buf.data() + (size_t)line * n,
                                  sizeof(uint32_t) * n, hipMemcpyHostToDevice, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
    }
    HIPCHK(hipStreamSynchronize(stream_));
    s.n_pts = n;
    bins_valid_ = false;
    invalidate_graphs();
    RC(push_scenes());
    return ICEMPM_OK;
}

int Engine::download(int scene, double *hssoa, unsigned capacity, int offset, int *n_out)
{
    RC(ensure_committed());
    SceneHost &s = scenes_[scene];
    const int n = s.n_pts;
    if (offset < 0 || (long long)offset + n > (long long)capacity)
        return fail(ICEMPM_ERR_CAPACITY, "transfer_from_device() HSSOA capacity");
    if (precision_ == ICEMPM_PRECISION_FP64) {
        for (int line = 0; line < kNumPtsLines && n > 0; line++)
            HIPCHK(hipMemcpyAsync(hssoa + (size_t)line * capacity + offset, (double *)pts_[0] + line * pitch_ + s.pt_off,
                                  sizeof(double) * n, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
    } else {
        std::vector<uint32_t> buf((size_t)kNumPtsLines * std::max(n, 1));
        for (int line = 0; line < kNumPtsLines && n > 0; line++)
            HIPCHK(hipMemcpyAsync(buf.data() + (size_t)line * n, (uint32_t *)pts_[0] + line * pitch_ + s.pt_off,
                                  sizeof(uint32_t) * n, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        const SceneDev &d = host_dev_[scene];
        const double inv = std::ldexp(1.0, -d.c.frac_bits);
        for (int i = 0; i < n; i++) {
            auto getf = [&](int line) {
                float f;
                std::memcpy(&f, &buf[(size_t)line * n + i], 4);
                return (double)f;
            };
            auto put = [&](int line, double v) { hssoa[(size_t)line * capacity + offset + i] = v; };
            const long long ud = (long long)buf[i];
            std::memcpy(hssoa + offset + i, &ud, 8);
            const int org[2] = {d.c.origin, d.c.origin_y};
            for (int k = 0; k < 2; k++)
                put(kLinePosX + k, ((double)buf[(size_t)(kLinePosX + k) * n + i] * inv - org[k]) * d.c.h);
            put(kLineP, getf(kLineP));
            put(kLineQ, getf(kLineQ));
            put(kLineJp, getf(kLineJp) + 1.0);
            put(kLineVelX, getf(kLineVelX));
            put(kLineVelX + 1, getf(kLineVelX + 1));
            put(kLineFe00, getf(kLineFe00) + 1.0);
            put(kLineFe00 + 1, getf(kLineFe00 + 1));
            put(kLineFe00 + 2, getf(kLineFe00 + 2));
            put(kLineFe00 + 3, getf(kLineFe00 + 3) + 1.0);
            for (int k = 0; k < 4; k++) put(kLineBp00 + k, getf(kLineBp00 + k));
        }
    }
    if (n_out) *n_out = n;
    return ICEMPM_OK;
}

// ---------------------------------------------------------------------------
// timing helpers
// ---------------------------------------------------------------------------
// During a graph capture an event record becomes a node of the graph only
// with the external flag (a plain record would only order the capture).
void Engine::record(hipEvent_t e)
{
    if (capturing_) hipEventRecordWithFlags(e, stream_, hipEventRecordExternal);
    else hipEventRecord(e, stream_);
}

void Engine::timed_begin(bool timed)
{
    if (!timed) return;
    // bounds the event pool of long single-phase loops (no synchronization
    // while a graph is being captured)
    if (ev_used_ >= 4096 && !capturing_) collect_timings();
    if (ev_used_ == ev_pool_.size()) {
        hipEvent_t a, b;
        hipEventCreate(&a);
        hipEventCreate(&b);
        ev_pool_.push_back({a, b});
        ev_cls_.push_back(0);
        ev_w_.push_back(1);
    }
    record(ev_pool_[ev_used_].first);
}

void Engine::timed_end(bool timed, int cls, int launches)
{
    if (!timed) return;
    record(ev_pool_[ev_used_].second);
    ev_cls_[ev_used_] = cls;
    ev_w_[ev_used_] = launches;
    ev_used_++;
}

int Engine::collect_timings()
{
    if (ev_used_ == 0) return ICEMPM_OK;
    HIPCHK(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < ev_used_; i++) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev_pool_[i].first, ev_pool_[i].second));
        cls_ms_[ev_cls_[i]] += ms;
        cls_cnt_[ev_cls_[i]] += ev_w_[i];
    }
    ev_used_ = 0;
    return ICEMPM_OK;
}

int Engine::get_kernel_stats(double *ms, long long *count, int n)
{
    RC(collect_timings());
    for (int k = 0; k < n && k < kNumCls; k++) {
        if (ms) ms[k] = cls_ms_[k];
        if (count) count[k] = cls_cnt_[k];
    }
    return ICEMPM_OK;
}

// ---------------------------------------------------------------------------
// phases
// ---------------------------------------------------------------------------
int Engine::zero_grid(int scene, int buf)
{
    const SceneHost &s = scenes_[scene];
    HIPCHK(launch_fill_zero(s.grid[buf], (size_t)s.grid_elems * kGridLines * sizeof(double), stream_));
    HIPCHK(launch_fill_zero(d_dirty_ + buf * n_segs_ + s.seg_off, sizeof(int) * (s.grid_elems >> kSegShift),
                            stream_));
    scenes_[scene].clean[buf] = true;
    return ICEMPM_OK;
}

#define LAUNCH_PREC(fn, ...) \
    (precision_ == ICEMPM_PRECISION_MIXED ? mixed::fn(__VA_ARGS__) : fp64::fn(__VA_ARGS__))

// The single-phase calls record hipEvents around their launch (the host-driven
// substep loop reads per-phase times like the reference's record_timings).
int Engine::reset_grid(int scene)
{
    RC(ensure_committed());
    timed_begin(true);
    const int rc = zero_grid(scene, 0);
    timed_end(true, kClsZero);
    return rc;
}

// Workgroups of a tile launch.  The kernels loop over the tile list, and a
// launch of exactly the workgroups the GPU holds at once (one generation) is
// the fastest: with more, the later generations start behind the first and the
// launch ends on a tail of partly busy CUs (4M-point sheet, fused launch: FP64
// 139.7 us with 8 workgroups per CU, 133.5 us with 2, its residency; mixed
// 106.0 us with 8, 103.6 us with 4, its residency, 133 us with 2).
int Engine::tile_blocks(int mode)
{
    int &r = resident_[mode & 7];
    if (r <= 0) {
        r = LAUNCH_PREC(tiles_resident_per_cu, args(), mode);
        if (r <= 0) r = 8;   // the runtime could not tell: several generations
    }
    return (int)std::min<long long>(std::max<long long>(max_tiles_, 1), (long long)n_cus_ * r);
}

int Engine::p2g(int scene)
{
    RC(ensure_committed());
    const StepArgs a = args();
    timed_begin(true);
    if (bins_valid_ && nscenes() == 1) {
        HIPCHK(LAUNCH_PREC(launch_tiles, a, kTileP2G, 0, 0, 0, 0, tile_blocks(kTileP2G), stream_));
    } else {
        HIPCHK(LAUNCH_PREC(launch_p2g_direct, a, scene, scenes_[scene].n_pts, 0, stream_));
    }
    timed_end(true, kClsP2G);
    scenes_[scene].clean[0] = false;
    return ICEMPM_OK;
}

int Engine::update_nodes(int scene, double t, double ind_x, double ind_y)
{
    RC(ensure_committed());
    SceneClock ck{};
    ck.t = t;
    ck.ind_x = ind_x;
    ck.ind_y = ind_y;
    ck.step = scenes_[scene].sim_step;
    ck.next_step = scenes_[scene].sim_step;
    HIPCHK(hipMemcpyAsync(d_clocks_ + clock_slot_ * nscenes() + scene, &ck, sizeof ck, hipMemcpyHostToDevice,
                          stream_));
    StepArgs a = args();
    // only this scene's chunks: the chunk table is scene-major
    const NodeChunk *first = nullptr;
    int nchunks = 0;
    {
        long long k = 0;
        for (int i = 0; i < nscenes(); i++) {
            const long long nseg = scenes_[i].grid_elems >> kSegShift;
            const int len = update_run_length(scenes_[i].grid_elems);
            const int nc = (int)((nseg + len - 1) / len);
            if (i == scene) { first = d_chunks_ + k; nchunks = nc; }
            k += nc;
        }
    }
    a.chunks = first;
    a.n_chunks = nchunks;
    timed_begin(true);
    HIPCHK(fp64::launch_update(a, 0, 0, 0, update_blocks(nchunks), stream_));
    timed_end(true, kClsUpdate);
    HIPCHK(hipStreamSynchronize(stream_));   // ck lives on this stack frame
    return ICEMPM_OK;
}

int Engine::g2p(int scene, int recordPQ)
{
    RC(ensure_committed());
    const StepArgs a = args();
    timed_begin(true);
    if (bins_valid_ && nscenes() == 1) {
        HIPCHK(LAUNCH_PREC(launch_tiles, a, kTileG2P, 0, 0, 0, recordPQ ? 1 : 0, tile_blocks(kTileG2P), stream_));
    } else {
        HIPCHK(LAUNCH_PREC(launch_g2p_direct, a, scene, scenes_[scene].n_pts, 0, recordPQ ? 1 : 0, stream_));
    }
    timed_end(true, kClsG2P);
    scenes_[scene].sim_step++;
    return ICEMPM_OK;
}

int Engine::reset_indenter_force()
{
    RC(ensure_committed());
    long long forces = 0;
    for (auto &s : scenes_) forces += 2LL * s.prms.n_indenter_subdivisions;
    HIPCHK(launch_fill_zero(ind_force_mem_, sizeof(double) * forces, stream_));
    return reset_kernel_stats();
}

int Engine::reset_kernel_stats()
{
    RC(collect_timings());
    for (int k = 0; k < kNumCls; k++) { cls_ms_[k] = 0; cls_cnt_[k] = 0; }
    return ICEMPM_OK;
}

int Engine::synchronize()
{
    HIPCHK(hipStreamSynchronize(stream_));
    return ICEMPM_OK;
}

int Engine::get_indenter_force(int scene, double *out, int n)
{
    if (n < 0) return fail(ICEMPM_ERR_ARGUMENT, "get_indenter_force: negative length");
    RC(ensure_committed());
    const SceneHost &s = scenes_[scene];
    const int m = std::min(n, 2 * s.prms.n_indenter_subdivisions);
    HIPCHK(hipMemcpyAsync(out, s.ind_force, sizeof(double) * m, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return ICEMPM_OK;
}

int Engine::get_error_code(uint8_t *code)
{
    RC(ensure_committed());
    std::vector<unsigned> e(nscenes());
    HIPCHK(hipMemcpyAsync(e.data(), d_err_, sizeof(unsigned) * nscenes(), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    unsigned m = 0;
    for (unsigned v : e) m = std::max(m, v);
    *code = (uint8_t)m;
    return ICEMPM_OK;
}

int Engine::get_position_map(int scene, int *fb, int *ox, int *oy)
{
    if (scene < 0 || scene >= nscenes() || !scenes_[scene].have_prms)
        return fail(ICEMPM_ERR_STATE, "get_position_map: parameters not set");
    SceneDev d;
    build_scene_dev(scene, d);
    *fb = d.c.frac_bits;
    *ox = d.c.origin;
    *oy = d.c.origin_y;
    return ICEMPM_OK;
}

int Engine::get_grid(int scene, int buf, double *out, long long n_nodes)
{
    RC(ensure_committed());
    if (buf < 0 || buf >= kGridBuffers) return fail(ICEMPM_ERR_ARGUMENT, "get_grid: buffer 0, 1 or 2");
    const SceneHost &s = scenes_[scene];
    const long long nodes = (long long)(s.prms.GridXTotal + 2 * s.prms.GridHaloSize) * s.prms.GridY;
    if (n_nodes != nodes) return fail(ICEMPM_ERR_ARGUMENT, "get_grid: node count mismatch");
    HIPCHK(hipStreamSynchronize(stream_));
    for (int line = 0; line < kGridLines; line++)
        HIPCHK(hipMemcpy(out + line * nodes, (double *)s.grid[buf] + line * s.grid_elems, sizeof(double) * nodes,
                         hipMemcpyDeviceToHost));
    return ICEMPM_OK;
}

// ---------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------
int Engine::build_tiles()
{
    const StepArgs a = args();
    HIPCHK(launch_build_tiles(a, host_dev_.data(), nscenes(), bins_, d_tiles_, d_n_tiles_, max_tiles_, stream_));
    int nt = 0;
    HIPCHK(hipMemcpy(&nt, d_n_tiles_, sizeof(int), hipMemcpyDeviceToHost));
    if (nt < 0 || nt > max_tiles_) {
        bins_valid_ = false;
        return fail(ICEMPM_ERR_STATE, "build_tiles: tile count out of range");
    }
    n_live_tiles_ = nt;
    return ICEMPM_OK;
}

// Captured chunk graphs survive a sort: the binned kernels read the tile
// table, cell runs and point counts from device memory, and the swapped point
// buffer is part of the graph key.  Only the unbinned launches take point
// counts by value, so a sort that drops disabled points clears the cache.
int Engine::sort()
{
    RC(ensure_committed());
    StepArgs a = args();
    timed_begin(false);
    HIPCHK(launch_sort(a, host_dev_.data(), nscenes(), bins_, elem_bytes(), pts_[0], pts_[1], stream_));
    std::swap(pts_[0], pts_[1]);
    pts_flip_ ^= 1;
    std::vector<int> cnt(nscenes());
    HIPCHK(hipMemcpyAsync(cnt.data(), bins_.scene_count, sizeof(int) * nscenes(), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    bool counts_changed = false;
    for (int i = 0; i < nscenes(); i++) {
        const int m = scenes_[i].n_pts > 0 ? cnt[i] : 0;
        counts_changed = counts_changed || m != scenes_[i].n_pts;
        scenes_[i].n_pts = m;
    }
    if (counts_changed) invalidate_graphs();
    RC(push_scenes());
    RC(build_tiles());
    bins_valid_ = true;
    return ICEMPM_OK;
}

int Engine::get_bins(int scene, int32_t *keys, int32_t *perm, int n, int32_t *cell_counts, int n_cells)
{
    if (n < 0 || n_cells < 0) return fail(ICEMPM_ERR_ARGUMENT, "get_bins: negative length");
    if (!bins_valid_) return fail(ICEMPM_ERR_STATE, "get_bins: no sort since the last upload");
    const SceneHost &s = scenes_[scene];
    const int m = std::min(n, s.n_pts);
    if (keys) HIPCHK(hipMemcpy(keys, bins_.sorted_keys + s.pt_off, sizeof(int) * m, hipMemcpyDeviceToHost));
    if (perm) {
        HIPCHK(hipMemcpy(perm, bins_.perm + s.pt_off, sizeof(int) * m, hipMemcpyDeviceToHost));
        for (int i = 0; i < m; i++) perm[i] -= (int32_t)s.pt_off;
    }
    if (cell_counts) {
        const int nc = (s.prms.GridXTotal + 2 * s.prms.GridHaloSize) * s.prms.GridY;
        std::vector<int> st(nc + 1);
        HIPCHK(hipMemcpy(st.data(), bins_.cell_start + s.cells_off, sizeof(int) * (nc + 1), hipMemcpyDeviceToHost));
        for (int i = 0; i < std::min(nc, n_cells); i++) cell_counts[i] = st[i + 1] - st[i];
    }
    return ICEMPM_OK;
}

// ---------------------------------------------------------------------------
// frame stepping
// ---------------------------------------------------------------------------
// A call of nsub substeps is a sequence of segments of one kind of unit:
//   binned, fused:   Head [P2G(b)], nsub-1 x Chain, Tail
//     Chain [grid update + G2P from grid b, P2G into b+1, clear b+2; the
//            points set aside, by its last workgroup], then b = b+1 (mod 3)
//     Tail  [grid update + G2P from grid b, clear b+2; the points set aside]
//   binned, unfused: nsub x Unfused [P2G(b), update(b), G2P(b)], b = b^1 (0, 1)
//   unbinned:        nsub x Direct [zero, P2G direct, update, G2P direct]
// Every unit but the head advances the clock slot.  The tile launch of a
// Chain or Tail unit forms the node velocities of grid b from its momenta (the
// grid update folded in), reading clock slot cs and writing the advanced
// clock to cs^1, which the later launches of the substep read.
int Engine::emit_units(int kind, int n, int b, int cs, bool timed_all, int sample, int k0, bool span)
{
    StepArgs a = args();
    SceneClock *slot[2] = {d_clocks_, d_clocks_ + nscenes()};
    const int nb_upd = update_blocks(n_chunks_);
    const bool span_chain = span && timed_all && kind == kUnitChain && n > 0;
    timed_begin(span_chain);
    for (int j = 0; j < n; j++) {
        const bool timed = !span_chain && timed_all && (sample <= 1 || (k0 + j) % sample == 0);
        StepArgs au = a;               // the substep's clock
        au.clocks = slot[cs];
        au.clocks_next = slot[cs ^ 1];
        StepArgs an = a;               // after the update: the advanced clock
        an.clocks = slot[cs ^ 1];
        an.clocks_next = slot[cs];
        switch (kind) {
        case kUnitHead:
            timed_begin(timed);
            HIPCHK(LAUNCH_PREC(launch_tiles, au, kTileP2G, b, b, b, -1, tile_blocks(kTileP2G), stream_));
            timed_end(timed, kClsP2G);
            continue;                  // part of substep 0: no clock or buffer change
        case kUnitChain:
        case kUnitTail: {
            const int mode = kind == kUnitChain ? kTileChain : kTileTail;
            const int w = kind == kUnitChain ? (b + 1) % kGridBuffers : b, z = (b + 2) % kGridBuffers;
            timed_begin(timed);
            HIPCHK(LAUNCH_PREC(launch_tiles, au, mode, b, w, z, -1, tile_blocks(mode), stream_));
            timed_end(timed, kind == kUnitChain ? kClsFused : kClsG2P);
            b = w;
            cs ^= 1;
            continue;
        }
        case kUnitDirect:
            timed_begin(timed);
            for (int i = 0; i < nscenes(); i++) {
                const SceneHost &s = scenes_[i];
                HIPCHK(launch_fill_zero(s.grid[0], (size_t)s.grid_elems * kGridLines * sizeof(double), stream_));
            }
            HIPCHK(launch_fill_zero(d_dirty_, sizeof(int) * n_segs_, stream_));
            timed_end(timed, kClsZero);
            timed_begin(timed);
            for (int i = 0; i < nscenes(); i++)
                HIPCHK(LAUNCH_PREC(launch_p2g_direct, au, i, scenes_[i].n_pts, 0, stream_));
            timed_end(timed, kClsP2G);
            timed_begin(timed);
            HIPCHK(fp64::launch_update(au, 0, 0, 1, nb_upd, stream_));
            timed_end(timed, kClsUpdate);
            timed_begin(timed);
            for (int i = 0; i < nscenes(); i++)
                HIPCHK(LAUNCH_PREC(launch_g2p_direct, an, i, scenes_[i].n_pts, 0, -1, stream_));
            timed_end(timed, kClsG2P);
            cs ^= 1;
            continue;
        default:   // kUnitUnfused
            timed_begin(timed);
            HIPCHK(LAUNCH_PREC(launch_tiles, au, kTileP2G, b, b, b, -1, tile_blocks(kTileP2G), stream_));
            timed_end(timed, kClsP2G);
            timed_begin(timed);
            HIPCHK(fp64::launch_update(au, b, 1, 1, nb_upd, stream_));
            timed_end(timed, kClsUpdate);
            timed_begin(timed);
            HIPCHK(LAUNCH_PREC(launch_tiles, an, kTileG2P, b, b, b, -1, tile_blocks(kTileG2P), stream_));
            timed_end(timed, kClsG2P);
            cs ^= 1;
            b ^= 1;
            continue;
        }
    }
    timed_end(span_chain, kClsFused, n);
    return ICEMPM_OK;
}

// Clock slot and grid buffer after n units of a kind (see emit_units).
static void advance_units(int kind, int n, int &b, int &cs)
{
    if (kind == Engine::kUnitHead) return;
    cs ^= n & 1;
    if (kind == Engine::kUnitChain) b = (b + n) % kGridBuffers;
    else if (kind == Engine::kUnitUnfused) b ^= n & 1;
}

// The graph of n units of a kind from (b, cs), captured on first use.  The
// cache holds chunks of 1, 2, 4, ..., kMaxChunk units, so its size is bounded
// whatever substep counts the callers use; a captured chunk refers to device
// buffers that only commit(), set_tile_shape() and upload() replace (they
// clear the cache), and to the point buffer pair member that the sort swaps
// (pts_flip_ is part of the key).
int Engine::chunk_graph(int kind, int n, int b, int cs, int tmode, bool first_timed, GraphExec *&out)
{
    const int sample = tmode == 2 ? kTimingSample : 1;
    const GraphKey key{kind, n, b, cs, pts_flip_, tmode, first_timed ? 1 : 0};
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
        // step() makes room before a call's first launch: clearing the cache
        // here would destroy graphs this call launched and still reads
        if (graphs_.size() >= kMaxGraphs + kMaxGraphsPerCall)
            return fail(ICEMPM_ERR_STATE, "graph cache over its bound within a call");
        // The capture records the timing events into the graph: it takes them
        // from the pool (emptied first) and keeps them.
        if (tmode) RC(collect_timings());
        // workgroup counts are runtime queries: make them before the capture
        for (int m : {(int)kTileG2P, (int)kTileP2G, (int)kTileTail, (int)kTileChain}) tile_blocks(m);
        HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        capturing_ = true;
        // first_timed with sampling: unit 0 of the chunk is a sampled substep
        // of its call, and so is every unit a multiple of the sample after it
        const int rc = emit_units(kind, n, b, cs, tmode != 0, sample, first_timed ? 0 : 1, tmode == 3);
        capturing_ = false;
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(stream_, &g);
        GraphExec gx;
        gx.ev.assign(ev_pool_.begin(), ev_pool_.begin() + ev_used_);
        gx.cls.assign(ev_cls_.begin(), ev_cls_.begin() + ev_used_);
        gx.w.assign(ev_w_.begin(), ev_w_.begin() + ev_used_);
        ev_pool_.erase(ev_pool_.begin(), ev_pool_.begin() + ev_used_);
        ev_cls_.erase(ev_cls_.begin(), ev_cls_.begin() + ev_used_);
        ev_w_.erase(ev_w_.begin(), ev_w_.begin() + ev_used_);
        ev_used_ = 0;
        auto drop = [&]() {
            for (auto &e : gx.ev) {
                hipEventDestroy(e.first);
                hipEventDestroy(e.second);
            }
        };
        if (rc != ICEMPM_OK) { if (g) hipGraphDestroy(g); drop(); return rc; }
        if (ec != hipSuccess) { drop(); return hipfail(ec, "hipStreamEndCapture"); }
        const hipError_t ei = hipGraphInstantiate(&gx.exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) { drop(); return hipfail(ei, "hipGraphInstantiate"); }
        it = graphs_.emplace(key, gx).first;
        n_captures_++;
    }
    out = &it->second;
    return ICEMPM_OK;
}

// Adds the event times of a replayed timed graph to the kernel statistics.
int Engine::read_graph_events(const GraphExec &gx)
{
    for (size_t i = 0; i < gx.ev.size(); i++) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, gx.ev[i].first, gx.ev[i].second));
        cls_ms_[gx.cls[i]] += ms;
        cls_cnt_[gx.cls[i]] += gx.w[i];
    }
    return ICEMPM_OK;
}

// Runs n units of a kind starting at substep k0 of the call, as chunk graph
// replays (use_graph) or direct launches; b and cs are advanced.
int Engine::run_units(int kind, int n, int &b, int &cs, bool use_graph, int tmode, int k0,
                      std::vector<GraphExec *> &replayed)
{
    const int sample = tmode == 2 ? kTimingSample : 1;
    if (!use_graph) {
        RC(emit_units(kind, n, b, cs, tmode != 0, sample, k0, tmode == 3));
        advance_units(kind, n, b, cs);
        return ICEMPM_OK;
    }
    int k = k0;
    while (n > 0) {
        // largest power of two <= n, at most kMaxChunk: a chunk of 32 or more
        // units then starts at a multiple of 32 within the segment
        int len = kMaxChunk;
        while (len > n) len >>= 1;
        const bool first_timed = tmode == 2 ? (k % kTimingSample) == 0 : tmode != 0;
        GraphExec *gx = nullptr;
        RC(chunk_graph(kind, len, b, cs, tmode, first_timed, gx));
        if (tmode && std::find(replayed.begin(), replayed.end(), gx) != replayed.end()) {
            // the graph's events still hold its previous replay in this call: read them first
            HIPCHK(hipStreamSynchronize(stream_));
            RC(read_graph_events(*gx));
            replayed.erase(std::find(replayed.begin(), replayed.end(), gx));
        }
        HIPCHK(hipGraphLaunch(gx->exec, stream_));
        if (tmode) replayed.push_back(gx);
        advance_units(kind, len, b, cs);
        n -= len;
        k += len;
    }
    return ICEMPM_OK;
}

// The first grid buffer of a call.  Fused: w0 with buffers w0 and w0+1 clean
// in every scene (the head accumulates into w0, the first chain launch into
// w0+1), buffers zeroed if no such pair exists.  Unfused: a clean buffer 0 or 1
// (the grid update of each substep clears the other).  Unbinned: 0.
int Engine::start_buffer(bool fused, bool direct, int &b)
{
    b = 0;
    if (direct) return ICEMPM_OK;
    auto all_clean = [&](int k) {
        for (auto &s : scenes_)
            if (!s.clean[k]) return false;
        return true;
    };
    if (fused) {
        for (int w = 0; w < kGridBuffers; w++)
            if (all_clean(w) && all_clean((w + 1) % kGridBuffers)) { b = w; return ICEMPM_OK; }
        for (int i = 0; i < nscenes(); i++)
            for (int k = 0; k < 2; k++)
                if (!scenes_[i].clean[k]) RC(zero_grid(i, k));
        return ICEMPM_OK;
    }
    if (all_clean(0)) return ICEMPM_OK;
    if (all_clean(1)) { b = 1; return ICEMPM_OK; }
    for (int i = 0; i < nscenes(); i++) RC(zero_grid(i, 0));
    return ICEMPM_OK;
}

// Clean grid buffers after a call of nsub substeps from buffer b (see emit_units).
void Engine::mark_buffers(bool fused, bool direct, int b, int nsub)
{
    for (auto &s : scenes_) {
        if (direct) {
            s.clean[0] = false;
        } else if (!fused) {
            const int last = b ^ ((nsub - 1) & 1);
            s.clean[last] = false;
            s.clean[last ^ 1] = true;
        } else {
            // only the buffer the tail read holds data; the others were cleared
            const int last = (b + nsub - 1) % kGridBuffers;
            for (int k = 0; k < kGridBuffers; k++) s.clean[k] = k != last;
        }
    }
}

// The units of a call of nsub substeps (see emit_units), without graphs.
int Engine::launch_substeps(int nsub, int start_buf, bool fused, bool direct, bool timed, int sample)
{
    std::vector<GraphExec *> none;
    const int tmode = timed ? (sample > 1 ? 2 : 1) : 0;
    int b = start_buf, cs = clock_slot_;
    if (direct) {
        RC(run_units(kUnitDirect, nsub, b, cs, false, tmode, 0, none));
    } else if (!fused) {
        RC(run_units(kUnitUnfused, nsub, b, cs, false, tmode, 0, none));
    } else {
        RC(run_units(kUnitHead, 1, b, cs, false, tmode, 0, none));
        RC(run_units(kUnitChain, nsub - 1, b, cs, false, tmode, 0, none));
        RC(run_units(kUnitTail, 1, b, cs, false, tmode, nsub - 1, none));
    }
    clock_slot_ = cs;
    return ICEMPM_OK;
}

int Engine::step(int nsub, int flags)
{
    RC(ensure_committed());
    if (nsub <= 0) return ICEMPM_OK;
    if (flags & ICEMPM_STEP_RESORT) RC(sort());
    const bool direct = (flags & ICEMPM_STEP_DIRECT_P2G) || !bins_valid_;
    const bool fused = !(flags & ICEMPM_STEP_UNFUSED);
    const bool timed = (flags & ICEMPM_STEP_TIMINGS) != 0;
    const bool use_graph = (flags & ICEMPM_STEP_GRAPH) != 0;
    // timing: 1 an event pair per launch, 2 the same for one substep in
    // kTimingSample, 3 one pair around each run of chain launches
    const int tmode = timed ? ((flags & ICEMPM_STEP_SAMPLED) ? 2 : (flags & ICEMPM_STEP_SPAN) ? 3 : 1) : 0;

    // A call captures at most kMaxGraphsPerCall chunk graphs; the cache is
    // cleared between calls (the last call's replays have completed: it
    // synchronised on ev_b_), never within one.
    if (use_graph && graphs_.size() + kMaxGraphsPerCall > kMaxGraphs) invalidate_graphs();
    int start_buf = 0;
    RC(start_buffer(fused, direct, start_buf));
    HIPCHK(hipEventRecord(ev_a_, stream_));
    std::vector<GraphExec *> replayed;
    int b = start_buf, cs = clock_slot_;
    if (direct) {
        RC(run_units(kUnitDirect, nsub, b, cs, use_graph, tmode, 0, replayed));
    } else if (!fused) {
        RC(run_units(kUnitUnfused, nsub, b, cs, use_graph, tmode, 0, replayed));
    } else {
        RC(run_units(kUnitHead, 1, b, cs, use_graph, tmode, 0, replayed));
        RC(run_units(kUnitChain, nsub - 1, b, cs, use_graph, tmode, 0, replayed));
        RC(run_units(kUnitTail, 1, b, cs, use_graph, tmode, nsub - 1, replayed));
    }
    clock_slot_ = cs;
    HIPCHK(hipEventRecord(ev_b_, stream_));
    HIPCHK(hipEventSynchronize(ev_b_));
    HIPCHK(hipEventElapsedTime(&last_step_ms_, ev_a_, ev_b_));
    last_step_ms_ /= nsub;
    if (timed) {
        // the replays have completed (ev_b_ above): read their events
        for (GraphExec *gx : replayed) RC(read_graph_events(*gx));
        RC(collect_timings());
    }

    mark_buffers(fused, direct, start_buf, nsub);
    for (auto &s : scenes_) {
        for (int k = 0; k < nsub; k++) s.sim_time += s.prms.InitialTimeStep;
        s.sim_step += nsub;
    }
    return ICEMPM_OK;
}

// ---------------------------------------------------------------------------
// autotuner: time the fused launches of the binned pipeline for each candidate
// tile shape, from a state that has drifted for half a frame since its sort
// (as at the middle of a frame, where points that left their tile window take
// the global path), keep the fastest and restore the state.
// ---------------------------------------------------------------------------
int Engine::autotune(int n_trials, float *out_ms, int n_out, int *chosen)
{
    RC(ensure_committed());
    if (!bins_valid_) RC(sort());
    // candidates: tile shape (tx, ty) x lane spread, static tile distribution.
    // The dynamic queues (set_dyn_tiles) are not tried: on MI355X their
    // device-scope atomics on one counter serialise and the launch takes
    // about twice as long at every shape (measured on the 4M-point sheet).
    // Lane spread 1 won at every shape on the 4M-point sheet, so the other maps
    // are tried at two shapes only; the small tiles are for sparse scenes.
    const int ncand = kNumTileCandidates;
    const size_t pts_bytes = (size_t)kNumPtsLines * pitch_ * elem_bytes();
    std::vector<SceneHost> saved = scenes_;
    const int tx0 = tx_, ty0 = ty_, ls0 = lane_spread_, dy0 = dyn_tiles_, margin0 = margin_;
    // the state at the call, restored at the end (the drift may set error bytes)
    std::vector<unsigned> err0(nscenes());
    HIPCHK(hipMemcpyAsync(err0.data(), d_err_, sizeof(unsigned) * nscenes(), hipMemcpyDeviceToHost, stream_));
    void *orig = nullptr;
    HIPCHK(hipMalloc(&orig, pts_bytes));
    HIPCHK(hipMemcpyAsync(orig, pts_[0], pts_bytes, hipMemcpyDeviceToDevice, stream_));
    // drift: half a frame of substeps with the current shape, no re-sort
    int drift = 0;
    for (auto &sc : scenes_) drift = std::max(drift, sc.prms.UpdateEveryNthStep / 2);
    drift = std::min(drift, 250);
    int rc = drift > 0 ? step(drift, ICEMPM_STEP_GRAPH) : ICEMPM_OK;
    // the drifted state, restored before each trial, sits in the sort buffer
    if (rc == ICEMPM_OK) {
        const hipError_t e = hipMemcpyAsync(pts_[1], pts_[0], pts_bytes, hipMemcpyDeviceToDevice, stream_);
        if (e != hipSuccess) rc = hipfail(e, "autotune: state copy");
    }
    std::vector<SceneHost> drifted = scenes_;
    float best = 1e30f;
    int best_i = -1, best_margin = margin0;
    n_trials = std::max(1, n_trials);
    // one trial: a warm-up substep, then n_trials fused units timed alone
    // (the head P2G and the tail G2P are outside the events); then the drifted
    // points, clocks and clean grids are restored.  ms = time per fused unit,
    // or -1 if the shape does not fit
    auto trial = [&](const int *cf, int margin, float &ms) -> int {
        lane_spread_ = cf[2];
        dyn_tiles_ = cf[3];
        ms = -1;
        if (set_tile_shape(cf[0], cf[1], margin) != ICEMPM_OK) return ICEMPM_OK;
        RC(step(1, 0));   // warm-up
        std::vector<GraphExec *> none;
        int b = 0, cs = clock_slot_;
        RC(start_buffer(true, false, b));
        RC(run_units(kUnitHead, 1, b, cs, false, 0, 0, none));
        HIPCHK(hipEventRecord(ev_a_, stream_));
        RC(run_units(kUnitChain, n_trials, b, cs, false, 0, 0, none));
        HIPCHK(hipEventRecord(ev_b_, stream_));
        RC(run_units(kUnitTail, 1, b, cs, false, 0, n_trials + 1, none));
        clock_slot_ = cs;
        HIPCHK(hipEventSynchronize(ev_b_));
        HIPCHK(hipEventElapsedTime(&ms, ev_a_, ev_b_));
        ms /= n_trials;
        HIPCHK(hipMemcpyAsync(pts_[0], pts_[1], pts_bytes, hipMemcpyDeviceToDevice, stream_));
        for (int i = 0; i < nscenes(); i++) {
            for (int k = 0; k < kGridBuffers; k++) RC(zero_grid(i, k));
            scenes_[i].sim_time = drifted[i].sim_time;
            scenes_[i].sim_step = drifted[i].sim_step;
        }
        return push_clocks();
    };
    // Every shape is timed with a window drift margin of 0 and of 1 (results
    // in that order).  No margin pays where the points stay in their cells
    // between sorts (smaller windows: less staging and flushing); the margin
    // where they drift.  The best shape differs between the two margins, so
    // both are swept in full.
    for (int mg = 0; mg < 2 && rc == ICEMPM_OK; mg++) {
        for (int c = 0; c < ncand && rc == ICEMPM_OK; c++) {
            float ms;
            rc = trial(kTileCandidates[c], mg, ms);
            const int k = mg * ncand + c;
            if (k < n_out && out_ms) out_ms[k] = ms;
            if (rc == ICEMPM_OK && ms >= 0 && ms < best) {
                best = ms;
                best_i = c;
                best_margin = mg;
            }
        }
    }
    // the state at the call
    {
        const hipError_t e = hipMemcpyAsync(pts_[0], orig, pts_bytes, hipMemcpyDeviceToDevice, stream_);
        const hipError_t e2 =
            hipMemcpyAsync(d_err_, err0.data(), sizeof(unsigned) * nscenes(), hipMemcpyHostToDevice, stream_);
        hipStreamSynchronize(stream_);
        hipFree(orig);
        if (rc == ICEMPM_OK && e != hipSuccess) rc = hipfail(e, "autotune: state restore");
        if (rc == ICEMPM_OK && e2 != hipSuccess) rc = hipfail(e2, "autotune: error byte restore");
    }
    for (int i = 0; i < nscenes(); i++) {
        scenes_[i].sim_time = saved[i].sim_time;
        scenes_[i].sim_step = saved[i].sim_step;
        for (int k = 0; k < kGridBuffers && rc == ICEMPM_OK; k++) rc = zero_grid(i, k);
    }
    if (best_i < 0 || rc != ICEMPM_OK) {
        lane_spread_ = ls0;
        dyn_tiles_ = dy0;
        RC(set_tile_shape(tx0, ty0, margin0));
    } else {
        lane_spread_ = kTileCandidates[best_i][2];
        dyn_tiles_ = kTileCandidates[best_i][3];
        RC(set_tile_shape(kTileCandidates[best_i][0], kTileCandidates[best_i][1], best_margin));
    }
    invalidate_graphs();   // captured launches hold the lane map and distribution by value
    RC(rc);
    RC(push_clocks());
    RC(reset_indenter_force());
    if (chosen) *chosen = best_i < 0 ? -1 : best_margin * ncand + best_i;
    HIPCHK(hipStreamSynchronize(stream_));
    return ICEMPM_OK;
}

int Engine::get_tile_config(int *tx, int *ty, int *margin, int *lane_spread) const
{
    if (tx) *tx = tx_;
    if (ty) *ty = ty_;
    if (margin) *margin = margin_;
    if (lane_spread) *lane_spread = lane_spread_;
    return ICEMPM_OK;
}

}  // namespace icempm
