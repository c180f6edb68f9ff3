// icempm_run.cpp -- command-line driver of the host layer, the counterpart of
// the reference's cli/main.cpp: parse a configuration file, load its raw
// points, and call icy::Model::Step until SimulationEndTime (or --frames).
//
//   icempm_run CONFIG.json [--points FILE.h5] [--frames N] [--precision fp64|mixed]
//              [--host-loop] [--no-resort] [--kernel-timings] [--log-dir DIR]
//              [--snapshot-dir DIR] [--resume SNAPSHOT.h5] [--dump-soa FILE] [--quiet]
//   icempm_run --convert-raw IN.bin OUT.h5    (flat binary -> LoadRawPoints HDF5 layout)
//   icempm_run --read-raw IN.h5 OUT.bin       (LoadRawPoints HDF5 layout -> flat binary)
//
// Flat raw-point binary: int64 n, float32 volume, float32 pad, then n doubles
// x, n doubles y, n uint64 llGrainIDs.  --dump-soa writes uint32 capacity,
// uint32 size, then the 16 HostSideSOA lines (capacity doubles each).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "model.h"
#include "snapshotmanager.h"

namespace {

struct RawPoints {
    std::vector<double> x, y;
    std::vector<uint64_t> g;
    float volume = 0;
};

RawPoints read_flat(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    int64_t n = 0;
    float vol = 0, pad = 0;
    f.read((char *)&n, 8);
    f.read((char *)&vol, 4);
    f.read((char *)&pad, 4);
    if (n <= 0 || n > (1LL << 31)) throw std::runtime_error("bad point count in " + path);
    RawPoints r;
    r.x.resize(n);
    r.y.resize(n);
    r.g.resize(n);
    r.volume = vol;
    f.read((char *)r.x.data(), 8 * n);
    f.read((char *)r.y.data(), 8 * n);
    f.read((char *)r.g.data(), 8 * n);
    if (!f) throw std::runtime_error("short file " + path);
    return r;
}

int usage()
{
    std::fprintf(stderr, "usage: icempm_run CONFIG.json [--points FILE.h5] [--frames N] [--precision fp64|mixed]\n"
                         "       [--host-loop] [--no-resort] [--kernel-timings] [--log-dir DIR] [--snapshot-dir DIR]\n"
                         "       [--resume SNAPSHOT.h5] [--dump-soa FILE] [--quiet] [--disable-every K]\n"
                         "       icempm_run --convert-raw IN.bin OUT.h5 | --read-raw IN.h5 OUT.bin\n");
    return 2;
}

}  // namespace

#ifdef ICEMPM_HAVE_HDF5
#include "h5io.h"
// the reads of SnapshotManager::LoadRawPoints, without bringing a model up
static int read_raw_to_flat(const char *in, const char *out)
{
    long long n = 0;
    if (h5io_raw_count(in, &n)) throw std::runtime_error(h5io_last_error());
    std::vector<double> x(n), y(n);
    std::vector<uint64_t> g(n);
    float vol = 0;
    if (h5io_read_raw(in, x.data(), y.data(), g.data(), n, &vol)) throw std::runtime_error(h5io_last_error());
    std::ofstream o(out, std::ios::binary);
    const int64_t nn = n;
    const float pad = 0;
    o.write((const char *)&nn, 8);
    o.write((const char *)&vol, 4);
    o.write((const char *)&pad, 4);
    o.write((const char *)x.data(), 8 * n);
    o.write((const char *)y.data(), 8 * n);
    o.write((const char *)g.data(), 8 * n);
    return o ? 0 : 1;
}
#endif

int main(int argc, char **argv)
{
    try {
        if (argc < 2) return usage();
        const std::string a1 = argv[1];
        if (a1 == "--convert-raw" || a1 == "--read-raw") {
            if (argc != 4) return usage();
#ifndef ICEMPM_HAVE_HDF5
            std::fprintf(stderr, "built without HDF5\n");
            return 1;
#else
            if (a1 == "--read-raw") return read_raw_to_flat(argv[2], argv[3]);
            RawPoints r = read_flat(argv[2]);
            icy::SnapshotManager::SaveRawPoints(argv[3], r.x.data(), r.y.data(), r.g.data(), (int)r.x.size(),
                                                r.volume);
            return 0;
#endif
        }

        std::string config = a1, points, log_dir = "logs", snapshot_dir, resume, dump;
        int frames = -1, disable_every = 0;
        icy::Model model;
        icy::SnapshotManager snapshot;
        snapshot.model = &model;
        model.Reset();
        for (int i = 2; i < argc; i++) {
            const std::string o = argv[i];
            auto next = [&]() -> std::string {
                if (i + 1 >= argc) throw std::runtime_error("missing value after " + o);
                return argv[++i];
            };
            if (o == "--points") points = next();
            else if (o == "--frames") frames = std::atoi(next().c_str());
            else if (o == "--precision") {
                const std::string p = next();
                if (p == "fp64") model.gpu.precision = ICEMPM_PRECISION_FP64;
                else if (p == "mixed") model.gpu.precision = ICEMPM_PRECISION_MIXED;
                else throw std::runtime_error("unknown precision " + p);
            } else if (o == "--host-loop") model.hostDrivenSubsteps = true;
            else if (o == "--no-resort") model.resortEveryFrame = false;
            else if (o == "--kernel-timings") model.recordKernelTimings = true;
            else if (o == "--log-dir") log_dir = next();
            else if (o == "--snapshot-dir") snapshot_dir = next();
            else if (o == "--resume") resume = next();
            else if (o == "--dump-soa") dump = next();
            else if (o == "--quiet") model.verbose = false;
            else if (o == "--disable-every") disable_every = std::atoi(next().c_str());
            else return usage();
        }
        model.logDirectory = log_dir;

        const std::string raw = model.prms.ParseFile(config);
        if (!resume.empty()) snapshot.ReadSnapshot(resume);
        else snapshot.LoadRawPoints(points.empty() ? raw : points);
        if (disable_every > 0) {
            // test hook: every k-th loaded point disabled (as a point that left
            // the domain would be), the active count adjusted, points re-uploaded
            HostSideSOA &S = model.gpu.hssoa;
            int active = 0;
            for (unsigned i = 0; i < S.size; i++) {
                if (i % (unsigned)disable_every == 0) S.setUtility(i, S.utility(i) | ICEMPM_STATUS_DISABLED);
                else active++;
            }
            model.prms.nPtsTotal = active;
            model.gpu.transfer_ponts_to_device();
        }

        bool more = true;
        for (int f = 0; more && (frames < 0 || f < frames); f++) more = model.Step();
        model.gpu.synchronize();
        if (!snapshot_dir.empty()) snapshot.SaveSnapshot(snapshot_dir);

        if (!dump.empty()) {
            HostSideSOA &S = model.gpu.hssoa;
            std::ofstream o(dump, std::ios::binary);
            o.write((const char *)&S.capacity, 4);
            o.write((const char *)&S.size, 4);
            o.write((const char *)S.host_buffer, sizeof(double) * (size_t)S.capacity * icy::SimParams::nPtsArrays);
            if (!o) throw std::runtime_error("writing " + dump);
        }
        int uploads = 0;
        for (const GPU_Partition &p : model.gpu.partitions) uploads += p.uploads;
        std::printf("done: step %d time %.6f points %u uploads %d device_rebalances %d\n", model.prms.SimulationStep,
                    model.prms.SimulationTime, model.gpu.hssoa.size, uploads, model.gpu.device_rebalances);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "icempm_run: %s\n", e.what());
        return 1;
    }
}
