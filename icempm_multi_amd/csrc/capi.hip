// capi.hip -- extern "C" entry points of include/icempm.h.
//
// A partition is an Engine holding one scene; a batch is an Engine holding
// many.  Errors: every call returns an ICEMPM_* status and leaves a message
// for icempm_last_error() on the calling thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "../../include/icempm.h"
#include "engine.h"
#include "icempm_host.h"

namespace icempm {
static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace icempm

using icempm::Engine;

struct icempm_partition {
    std::unique_ptr<Engine> eng;
    int partition_id = 0;
    int cap = 0;
    int grid_x_cap = 0;
    bool allocated = false;
};

struct icempm_batch {
    std::unique_ptr<Engine> eng;
    std::vector<std::vector<double>> staged;   // host copies until a commit succeeds
    std::vector<int> staged_n;
    bool committed = false;
};

namespace {

int bad(const char *msg)
{
    icempm::set_last_error(msg);
    return ICEMPM_ERR_ARGUMENT;
}

int check_prms(const icempm_simparams *prms, size_t bytes)
{
    if (!prms) return bad("null parameters");
    if (bytes != sizeof(icempm_simparams)) return bad("SimParams size mismatch");
    return ICEMPM_OK;
}

template <class F>
int guarded(F &&f)
{
    try {
        return f();
    } catch (const std::bad_alloc &) {
        icempm::set_last_error("host allocation failed");
        return ICEMPM_ERR_NOMEM;
    } catch (const std::exception &e) {
        icempm::set_last_error(e.what());
        return ICEMPM_ERR_STATE;
    }
}

}  // namespace

extern "C" {

const char *icempm_version(void) { return "icempm-mi355x 0.1 (gfx950)"; }

const char *icempm_last_error(void) { return icempm::g_last_error.c_str(); }

int icempm_params_reset(icempm_simparams *prms)
{
    if (!prms) return bad("null parameters");
    icempm::params_reset(prms);
    return ICEMPM_OK;
}

int icempm_params_parse_json(icempm_simparams *prms, const char *json_text, char *input_raw_points, size_t n)
{
    if (!prms || !json_text) return bad("null argument");
    return guarded([&] {
        std::string raw, err;
        const int rc = icempm::params_parse_json(prms, json_text, raw, err);
        if (rc != ICEMPM_OK) {
            icempm::set_last_error(err);
            return rc;
        }
        if (input_raw_points && n > 0) {
            std::strncpy(input_raw_points, raw.c_str(), n - 1);
            input_raw_points[n - 1] = 0;
        }
        return ICEMPM_OK;
    });
}

int icempm_params_derive(icempm_simparams *prms)
{
    if (!prms) return bad("null parameters");
    icempm::params_camclay(prms);
    icempm::params_helper(prms);
    return ICEMPM_OK;
}

int icempm_prepare_raw_points(icempm_simparams *prms, double *hssoa, unsigned hssoa_capacity, int n_points,
                              float volume, int *n_kept)
{
    if (!prms || !hssoa) return bad("null argument");
    return guarded([&] {
        std::string err;
        const int m = icempm::prepare_raw_points(prms, hssoa, hssoa_capacity, n_points, volume, err);
        if (m < 0) {
            icempm::set_last_error(err);
            return m;
        }
        if (n_kept) *n_kept = m;
        return ICEMPM_OK;
    });
}

// ---------------------------------------------------------------- partition

int icempm_partition_create(int device, int partition_id, int precision, icempm_partition **out)
{
    if (!out) return bad("null output");
    if (precision != ICEMPM_PRECISION_FP64 && precision != ICEMPM_PRECISION_MIXED) return bad("unknown precision");
    return guarded([&] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) {
            icempm::set_last_error("No available HIP devices");
            return ICEMPM_ERR_HIP;
        }
        if (device < 0 || device >= n) return bad("device index out of range");
        auto *p = new icempm_partition;
        p->eng.reset(new Engine(device, precision));
        p->partition_id = partition_id;
        *out = p;
        return ICEMPM_OK;
    });
}

int icempm_partition_destroy(icempm_partition *p)
{
    delete p;
    return ICEMPM_OK;
}

int icempm_partition_allocate(icempm_partition *p, int n_points_capacity, int grid_x_capacity)
{
    if (!p) return bad("null partition");
    if (n_points_capacity <= 0) return bad("allocate: capacity must be positive");
    if (p->allocated || p->eng->nscenes() != 0) return bad("allocate: already allocated");
    return guarded([&] {
        const int id = p->eng->add_scene(nullptr, n_points_capacity);
        if (id < 0) return id;
        p->cap = n_points_capacity;
        p->grid_x_cap = grid_x_capacity;
        p->allocated = true;
        return ICEMPM_OK;
    });
}

int icempm_partition_update_constants(icempm_partition *p, const icempm_simparams *prms, size_t prms_bytes)
{
    if (!p) return bad("null partition");
    int rc = check_prms(prms, prms_bytes);
    if (rc) return rc;
    if (!p->allocated) {
        icempm::set_last_error("update_constants: allocate first");
        return ICEMPM_ERR_STATE;
    }
    return guarded([&] {
        int r = p->eng->set_params(0, *prms);
        if (r) return r;
        return p->eng->ensure_committed();
    });
}

int icempm_partition_set_grid_span(icempm_partition *p, int gridx_offset, int gridx_partition)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->set_grid_span(0, gridx_offset, gridx_partition); });
}

int icempm_partition_transfer_points_from_soa_to_device(icempm_partition *p, const double *hssoa,
                                                        unsigned hssoa_capacity, int point_idx_offset, int n_points)
{
    if (!p || !hssoa) return bad("null argument");
    return guarded([&] { return p->eng->upload(0, hssoa, hssoa_capacity, point_idx_offset, n_points); });
}

int icempm_partition_transfer_from_device(icempm_partition *p, double *hssoa, unsigned hssoa_capacity,
                                          int point_idx_offset)
{
    if (!p || !hssoa) return bad("null argument");
    return guarded([&] {
        int n = 0;
        return p->eng->download(0, hssoa, hssoa_capacity, point_idx_offset, &n);
    });
}

int icempm_partition_reset_grid(icempm_partition *p)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->reset_grid(0); });
}

int icempm_partition_reset_indenter_force_accumulator(icempm_partition *p)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->reset_indenter_force(); });
}

int icempm_partition_p2g(icempm_partition *p)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->p2g(0); });
}

int icempm_partition_update_nodes(icempm_partition *p, double simulation_time, double indenter_x, double indenter_y)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->update_nodes(0, simulation_time, indenter_x, indenter_y); });
}

int icempm_partition_g2p(icempm_partition *p, int recordPQ)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->g2p(0, recordPQ); });
}

int icempm_partition_synchronize(icempm_partition *p)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->synchronize(); });
}

int icempm_partition_get_indenter_force(icempm_partition *p, double *out, int n_doubles)
{
    if (!p || !out) return bad("null argument");
    return guarded([&] { return p->eng->get_indenter_force(0, out, n_doubles); });
}

int icempm_partition_get_error_code(icempm_partition *p, uint8_t *code)
{
    if (!p || !code) return bad("null argument");
    return guarded([&] { return p->eng->get_error_code(code); });
}

int icempm_partition_get_point_count(icempm_partition *p, int *n_points, int *n_disabled)
{
    if (!p) return bad("null partition");
    if (n_points) *n_points = p->eng->nscenes() ? p->eng->scene(0).n_pts : 0;
    if (n_disabled) *n_disabled = 0;
    return ICEMPM_OK;
}

int icempm_partition_get_timings(icempm_partition *p, float *out, int n)
{
    if (!p || !out) return bad("null argument");
    return guarded([&] {
        double ms[icempm::kNumCls];
        long long cnt[icempm::kNumCls];
        int rc = p->eng->get_kernel_stats(ms, cnt, icempm::kNumCls);
        if (rc) return rc;
        float t[8] = {0};
        t[0] = (float)ms[icempm::kClsP2G];
        t[2] = (float)ms[icempm::kClsUpdate];
        // the binned stepping folds the grid update into its G2P launches
        t[3] = (float)(ms[icempm::kClsG2P] + ms[icempm::kClsFused]);
        for (int k = 0; k < icempm::kNumCls; k++)
            if (k != icempm::kClsSort) t[6] += (float)ms[k];
        t[7] = (float)ms[icempm::kClsSort];
        for (int k = 0; k < n && k < 8; k++) out[k] = t[k];
        return ICEMPM_OK;
    });
}

int icempm_partition_get_kernel_stats(icempm_partition *p, double *ms, long long *count, int n)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->get_kernel_stats(ms, count, n); });
}

int icempm_partition_step(icempm_partition *p, int n_substeps, int flags)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->step(n_substeps, flags); });
}

int icempm_partition_get_clock(icempm_partition *p, double *simulation_time, int *simulation_step)
{
    if (!p || !simulation_time || !simulation_step) return bad("null argument");
    if (p->eng->nscenes() == 0) return bad("not allocated");
    p->eng->clock(0, simulation_time, simulation_step);
    return ICEMPM_OK;
}

int icempm_partition_sort_points(icempm_partition *p)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->sort(); });
}

int icempm_partition_get_bins(icempm_partition *p, int32_t *keys, int32_t *perm, int n, int32_t *cell_counts,
                              int n_cells)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->get_bins(0, keys, perm, n, cell_counts, n_cells); });
}

int icempm_partition_get_grid(icempm_partition *p, int buf, double *out, long long n_nodes)
{
    if (!p || !out) return bad("null argument");
    return guarded([&] { return p->eng->get_grid(0, buf, out, n_nodes); });
}

int icempm_partition_get_position_map(icempm_partition *p, int *frac_bits, int *origin_x, int *origin_y)
{
    if (!p || !frac_bits || !origin_x || !origin_y) return bad("null argument");
    return guarded([&] { return p->eng->get_position_map(0, frac_bits, origin_x, origin_y); });
}

int icempm_partition_set_tile_shape(icempm_partition *p, int tx, int ty, int margin)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->set_tile_shape(tx, ty, margin); });
}

int icempm_partition_set_lane_map(icempm_partition *p, int lane_spread)
{
    if (!p) return bad("null partition");
    if (lane_spread < 0 || lane_spread > 2) return bad("lane_spread must be 0, 1 or 2");
    return guarded([&] {
        p->eng->set_lane_spread(lane_spread);
        return ICEMPM_OK;
    });
}

int icempm_partition_set_tile_schedule(icempm_partition *p, int dynamic)
{
    if (!p) return bad("null partition");
    if (dynamic != 0 && dynamic != 1) return bad("dynamic must be 0 or 1");
    return guarded([&] {
        p->eng->set_dyn_tiles(dynamic);
        return ICEMPM_OK;
    });
}

int icempm_tile_candidates(int *tx, int *ty, int *lane_spread, int n)
{
    for (int k = 0; k < n && k < icempm::kNumTileCandidates; k++) {
        if (tx) tx[k] = icempm::kTileCandidates[k][0];
        if (ty) ty[k] = icempm::kTileCandidates[k][1];
        if (lane_spread) lane_spread[k] = icempm::kTileCandidates[k][2];
    }
    return icempm::kNumTileCandidates;
}

int icempm_partition_autotune(icempm_partition *p, int n_trials, float *out_ms, int n_out, int *chosen)
{
    if (!p) return bad("null partition");
    return guarded([&] { return p->eng->autotune(n_trials, out_ms, n_out, chosen); });
}

int icempm_partition_get_tile_config(icempm_partition *p, int *tx, int *ty, int *margin, int *lane_spread)
{
    if (!p) return bad("null partition");
    return p->eng->get_tile_config(tx, ty, margin, lane_spread);
}

int icempm_partition_get_graph_stats(icempm_partition *p, long long *captures, int *cached)
{
    if (!p) return bad("null partition");
    if (captures) *captures = p->eng->graph_captures();
    if (cached) *cached = p->eng->cached_graphs();
    return ICEMPM_OK;
}

// ---------------------------------------------------------------- batch

int icempm_batch_create(int device, int precision, icempm_batch **out)
{
    if (!out) return bad("null output");
    if (precision != ICEMPM_PRECISION_FP64 && precision != ICEMPM_PRECISION_MIXED) return bad("unknown precision");
    return guarded([&] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) {
            icempm::set_last_error("No available HIP devices");
            return ICEMPM_ERR_HIP;
        }
        if (device < 0 || device >= n) return bad("device index out of range");
        auto *b = new icempm_batch;
        b->eng.reset(new Engine(device, precision));
        *out = b;
        return ICEMPM_OK;
    });
}

int icempm_batch_destroy(icempm_batch *b)
{
    delete b;
    return ICEMPM_OK;
}

int icempm_batch_add_scene(icempm_batch *b, const icempm_simparams *prms, size_t prms_bytes, const double *hssoa,
                           unsigned hssoa_capacity, int n_points, int *scene_id)
{
    if (!b || !hssoa) return bad("null argument");
    int rc = check_prms(prms, prms_bytes);
    if (rc) return rc;
    if (b->committed) {
        icempm::set_last_error("add_scene after commit");
        return ICEMPM_ERR_STATE;
    }
    if (n_points < 0 || (unsigned)n_points > hssoa_capacity) return bad("add_scene: bad point count");
    return guarded([&] {
        const int id = b->eng->add_scene(prms, std::max(n_points, 1));   // an empty scene keeps a 1-point range
        if (id < 0) return id;
        std::vector<double> copy((size_t)ICEMPM_N_PTS_ARRAYS * n_points);
        for (int line = 0; line < ICEMPM_N_PTS_ARRAYS; line++)
            std::memcpy(copy.data() + (size_t)line * n_points, hssoa + (size_t)line * hssoa_capacity,
                        sizeof(double) * n_points);
        b->staged.push_back(std::move(copy));
        b->staged_n.push_back(n_points);
        if (scene_id) *scene_id = id;
        return ICEMPM_OK;
    });
}

int icempm_batch_commit(icempm_batch *b)
{
    if (!b) return bad("null batch");
    if (b->committed) {
        icempm::set_last_error("batch already committed");
        return ICEMPM_ERR_STATE;
    }
    return guarded([&] {
        // the host copies are released only once every scene is on the device
        // and binned, so a failed commit can be retried
        int rc = b->eng->commit();
        if (rc) return rc;
        for (int i = 0; i < (int)b->staged.size(); i++) {
            rc = b->eng->upload(i, b->staged[i].data(), (unsigned)b->staged_n[i], 0, b->staged_n[i]);
            if (rc) return rc;
        }
        rc = b->eng->sort();
        if (rc) return rc;
        b->staged.clear();
        b->staged.shrink_to_fit();
        b->staged_n.clear();
        b->committed = true;
        return ICEMPM_OK;
    });
}

int icempm_batch_step(icempm_batch *b, int n_substeps, int flags)
{
    if (!b) return bad("null batch");
    if (!b->committed) {
        icempm::set_last_error("batch not committed");
        return ICEMPM_ERR_STATE;
    }
    return guarded([&] { return b->eng->step(n_substeps, flags); });
}

int icempm_batch_transfer_from_device(icempm_batch *b, int scene_id, double *hssoa, unsigned hssoa_capacity)
{
    if (!b || !hssoa) return bad("null argument");
    if (scene_id < 0 || scene_id >= b->eng->nscenes()) return bad("bad scene id");
    return guarded([&] {
        int n = 0;
        return b->eng->download(scene_id, hssoa, hssoa_capacity, 0, &n);
    });
}

int icempm_batch_get_indenter_force(icempm_batch *b, int scene_id, double *out, int n_doubles)
{
    if (!b || !out) return bad("null argument");
    if (scene_id < 0 || scene_id >= b->eng->nscenes()) return bad("bad scene id");
    return guarded([&] { return b->eng->get_indenter_force(scene_id, out, n_doubles); });
}

int icempm_batch_get_scene_point_count(icempm_batch *b, int scene_id, int *n_points)
{
    if (!b || !n_points) return bad("null argument");
    if (scene_id < 0 || scene_id >= b->eng->nscenes()) return bad("bad scene id");
    *n_points = b->eng->scene(scene_id).n_pts;
    return ICEMPM_OK;
}

int icempm_batch_get_bins(icempm_batch *b, int scene_id, int32_t *keys, int32_t *perm, int n, int32_t *cell_counts,
                          int n_cells)
{
    if (!b) return bad("null batch");
    if (scene_id < 0 || scene_id >= b->eng->nscenes()) return bad("bad scene id");
    return guarded([&] { return b->eng->get_bins(scene_id, keys, perm, n, cell_counts, n_cells); });
}

int icempm_batch_get_grid(icempm_batch *b, int scene_id, int buf, double *out, long long n_nodes)
{
    if (!b || !out) return bad("null argument");
    if (scene_id < 0 || scene_id >= b->eng->nscenes()) return bad("bad scene id");
    return guarded([&] { return b->eng->get_grid(scene_id, buf, out, n_nodes); });
}

int icempm_batch_get_position_map(icempm_batch *b, int scene_id, int *frac_bits, int *origin_x, int *origin_y)
{
    if (!b || !frac_bits || !origin_x || !origin_y) return bad("null argument");
    if (scene_id < 0 || scene_id >= b->eng->nscenes()) return bad("bad scene id");
    return guarded([&] { return b->eng->get_position_map(scene_id, frac_bits, origin_x, origin_y); });
}

int icempm_batch_get_error_code(icempm_batch *b, uint8_t *code)
{
    if (!b || !code) return bad("null argument");
    return guarded([&] { return b->eng->get_error_code(code); });
}

int icempm_batch_get_scene_count(icempm_batch *b, int *n_scenes, long long *n_points_total)
{
    if (!b) return bad("null batch");
    if (n_scenes) *n_scenes = b->eng->nscenes();
    if (n_points_total) *n_points_total = b->eng->total_points();
    return ICEMPM_OK;
}

int icempm_batch_autotune(icempm_batch *b, int n_trials, float *out_ms, int n_out, int *chosen)
{
    if (!b) return bad("null batch");
    if (!b->committed) {
        icempm::set_last_error("batch not committed");
        return ICEMPM_ERR_STATE;
    }
    return guarded([&] { return b->eng->autotune(n_trials, out_ms, n_out, chosen); });
}

int icempm_batch_set_tile_shape(icempm_batch *b, int tx, int ty, int margin)
{
    if (!b) return bad("null batch");
    return guarded([&] { return b->eng->set_tile_shape(tx, ty, margin); });
}

int icempm_batch_get_tile_config(icempm_batch *b, int *tx, int *ty, int *margin, int *lane_spread)
{
    if (!b) return bad("null batch");
    return b->eng->get_tile_config(tx, ty, margin, lane_spread);
}

int icempm_batch_get_graph_stats(icempm_batch *b, long long *captures, int *cached)
{
    if (!b) return bad("null batch");
    if (captures) *captures = b->eng->graph_captures();
    if (cached) *cached = b->eng->cached_graphs();
    return ICEMPM_OK;
}

int icempm_batch_last_step_ms(icempm_batch *b, float *ms_per_substep)
{
    if (!b || !ms_per_substep) return bad("null argument");
    *ms_per_substep = b->eng->last_step_ms();
    return ICEMPM_OK;
}

int icempm_batch_get_kernel_stats(icempm_batch *b, double *ms, long long *count, int n)
{
    if (!b) return bad("null batch");
    return guarded([&] { return b->eng->get_kernel_stats(ms, count, n); });
}

int icempm_batch_reset_kernel_stats(icempm_batch *b)
{
    if (!b) return bad("null batch");
    return guarded([&] { return b->eng->reset_kernel_stats(); });
}

}  // extern "C"
