// scene_gen.cpp -- Poisson-disk point clouds for the synthetic scenes.
//
// The reference's generator fills a block with thinks::PoissonDiskSampling of
// radius sqrt(0.656 * area / n) (generator/2d/grainprocessor2d.cpp:40-56,
// 250-262; generator/poisson_disk_sampling/poisson_disk_sampling.h).  This is
// Bridson's published algorithm ("Fast Poisson disk sampling in arbitrary
// dimensions", SIGGRAPH 2007 sketches): a background grid of cell side
// r/sqrt(2) holds at most one sample per cell; an active sample draws up to 30
// candidates uniformly from the annulus [r, 2r) around it, keeps the first
// one that lies in the box at distance >= r from every sample, and leaves the
// active list when all 30 fail.  The newest active sample is expanded first
// (the paper picks one at random; this keeps the growing front in cache).
// The cloud fills about 0.63 * area / r^2 points.  Random numbers come from
// SplitMix64, so a seed fixes the cloud.
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/icempm.h"
#include "icempm_host.h"

namespace {

int fail(int code, const char *msg)
{
    icempm::set_last_error(msg);
    return code;
}

struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)
};

constexpr int kCandidates = 30;

}  // namespace

extern "C" int icempm_poisson_disk(double r, double x0, double y0, double x1, double y1, uint64_t seed, double *x,
                                   double *y, long long cap, long long *n_out)
{
    if (!n_out || (cap > 0 && (!x || !y))) return fail(ICEMPM_ERR_ARGUMENT, "poisson_disk: null argument");
    *n_out = 0;
    if (!(r > 0) || !(x1 > x0) || !(y1 > y0) || !std::isfinite(r) || !std::isfinite(x1 - x0) ||
        !std::isfinite(y1 - y0))
        return fail(ICEMPM_ERR_ARGUMENT, "poisson_disk: need r > 0 and a non-empty finite box");
    const double cs = r / std::sqrt(2.0);
    const double gw = std::ceil((x1 - x0) / cs), gh = std::ceil((y1 - y0) / cs);
    if (gw * gh > 2.0e9) return fail(ICEMPM_ERR_ARGUMENT, "poisson_disk: box too large for r");
    const long long nx = (long long)gw, ny = (long long)gh;
    // cell (cx, cy) at (cy + 2) * pitch + cx + 2: two empty cells of border on
    // every side, so the neighbourhood scan needs no bounds checks
    const long long pitch = nx + 4;
    std::vector<int> cell((size_t)(pitch * (ny + 4)), -1);
    std::vector<double> px, py;
    std::vector<int> active;
    SplitMix64 rng{seed};
    const double r2 = r * r;

    auto cell_at = [&](double vx, double vy) -> long long {
        const long long cx = std::min<long long>(nx - 1, (long long)((vx - x0) / cs));
        const long long cy = std::min<long long>(ny - 1, (long long)((vy - y0) / cs));
        return (cy + 2) * pitch + cx + 2;
    };
    auto add = [&](double vx, double vy) -> bool {
        if ((long long)px.size() >= cap) return false;
        cell[(size_t)cell_at(vx, vy)] = (int)px.size();
        active.push_back((int)px.size());
        px.push_back(vx);
        py.push_back(vy);
        return true;
    };
    // far enough from every sample: a sample within r lies at most 2 cells
    // away, and not in a corner cell of the 5 x 5 block (those are >= r away)
    // (nearest cells first: a sample in the candidate's own cell is always
    // within r, and most candidates fail there)
    static const int kOrder[5] = {0, -1, 1, -2, 2};
    auto free_at = [&](double vx, double vy) {
        const long long c0 = cell_at(vx, vy);
        for (int a = 0; a < 5; a++) {
            const int dj = kOrder[a];
            const int w = (dj == -2 || dj == 2) ? 3 : 5;
            const int *row = cell.data() + c0 + dj * pitch;
            for (int b = 0; b < w; b++) {
                const int k = row[kOrder[b]];
                if (k < 0) continue;
                const double dx = px[k] - vx, dy = py[k] - vy;
                if (dx * dx + dy * dy < r2) return false;
            }
        }
        return true;
    };

    if (!add(x0 + rng.uniform() * (x1 - x0), y0 + rng.uniform() * (y1 - y0)))
        return fail(ICEMPM_ERR_CAPACITY, "poisson_disk: capacity exceeded");
    while (!active.empty()) {
        const size_t ai = active.size() - 1;   // the newest active sample: keeps the front local
        const int a = active[ai];
        bool placed = false;
        for (int t = 0; t < kCandidates; t++) {
            // uniform in the annulus [r, 2r): rejection from the enclosing square
            double dx, dy, d2;
            do {
                dx = (4.0 * rng.uniform() - 2.0) * r;
                dy = (4.0 * rng.uniform() - 2.0) * r;
                d2 = dx * dx + dy * dy;
            } while (d2 < r2 || d2 >= 4.0 * r2);
            const double vx = px[a] + dx, vy = py[a] + dy;
            if (vx < x0 || vx >= x1 || vy < y0 || vy >= y1) continue;
            if (!free_at(vx, vy)) continue;
            if (!add(vx, vy)) return fail(ICEMPM_ERR_CAPACITY, "poisson_disk: capacity exceeded");
            placed = true;
            break;
        }
        if (!placed) active.pop_back();
    }
    for (size_t i = 0; i < px.size(); i++) {
        x[i] = px[i];
        y[i] = py[i];
    }
    *n_out = (long long)px.size();
    return ICEMPM_OK;
}
