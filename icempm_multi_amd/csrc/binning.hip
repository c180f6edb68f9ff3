// binning.hip -- on-device HostSideSOA::RemoveDisabledAndSort
// (simulation/hssoa/host_side_soa.cpp:50-61) and the tile tables of the
// binned kernels.
//
// Counting sort by the reference's cell key (proxypoint2d.cpp:164-169):
//   key = (int)(x/h + 0.5)*GridY + (int)(y/h + 0.5)
// 1. keys + per-cell histogram (int atomics), disabled points dropped;
// 2. exclusive scan of the concatenated per-scene histograms -> cell_start;
// 3. scatter of point indices into their cells (atomic cursors);
// 4. each cell's indices sorted ascending, so ties keep their previous order
//    and the result is the stable order, independent of atomic timing;
// 5. gather of the 16 SoA lines into the second point buffer.
// Built with -ffp-contract=off so that x*hinv+0.5 is rounded twice, as on the host.
#include <hip/hip_runtime.h>

#include <vector>

#include "kernels.h"

namespace icempm {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int scene_of_point(const SceneDev *sc, int nscenes, long long i)
{
    int lo = 0, hi = nscenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sc[mid].pt_off <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int scene_of_cell(const SceneDev *sc, int nscenes, long long j)
{
    int lo = 0, hi = nscenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sc[mid].cell_off <= j) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// (int)(x*hinv + 0.5), reference rounding (two roundings, C truncation)
__device__ __forceinline__ int cell_index_fp64(double x, double hinv) { return (int)(x * hinv + 0.5); }

// the same on a fixed-point coordinate x/h = q*2^-fb - origin, exactly
__device__ __forceinline__ int cell_index_fixed(unsigned q, int fb, int origin)
{
    const long long v = (long long)q + (1LL << (fb - 1)) - ((long long)origin << fb);
    return v >= 0 ? (int)(v >> fb) : -(int)((-v) >> fb);
}

template <int EB>
__global__ void __launch_bounds__(kBlock) k_keys(StepArgs a, BinWork w, const void *src)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= w.n_pts_total) return;
    const int s = scene_of_point(a.scenes, a.nscenes, i);
    const SceneDev &S = a.scenes[s];
    const long long local = i - S.pt_off;
    if (local < 0 || local >= S.n_pts) { w.keys[i] = -2; return; }
    int ix, iy;
    if (EB == 8) {
        const double *P = (const double *)src;
        const long long ud = __double_as_longlong(P[i]);
        if (ud & kDisabled) { w.keys[i] = -1; return; }
        ix = cell_index_fp64(P[kLinePosX * a.pitch + i], S.c.hinv);
        iy = cell_index_fp64(P[(kLinePosX + 1) * a.pitch + i], S.c.hinv);
    } else {
        const unsigned *P = (const unsigned *)src;
        if (P[i] & (unsigned)kDisabled) { w.keys[i] = -1; return; }
        ix = cell_index_fixed(P[kLinePosX * a.pitch + i], S.c.frac_bits, S.c.origin);
        iy = cell_index_fixed(P[(kLinePosX + 1) * a.pitch + i], S.c.frac_bits, S.c.origin_y);
    }
    const int col = min(max(ix - S.c.xoff + S.c.halo, 0), S.c.ncols - 1);
    const int row = min(max(iy, 0), S.c.gridY - 1);
    const int bin = col * S.c.gridY + row;
    w.keys[i] = bin;
    atomicAdd(&w.cell_count[S.cell_off + bin], 1);
}

// Block-level exclusive scan; block totals to `sums` (may be null).
__global__ void __launch_bounds__(1024) k_scan_block(const int *in, int *out, int *sums, long long n)
{
    __shared__ int sh[1024];
    const long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    const int v = i < n ? in[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = sh[threadIdx.x] - v;
    if (sums && threadIdx.x == 1023) sums[blockIdx.x] = sh[1023];
}

__global__ void __launch_bounds__(1024) k_scan_add(int *out, const int *sums, long long n)
{
    const long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    if (i < n) out[i] += sums[blockIdx.x];
}

// cell_start = (global exclusive scan) - (scan at the scene's first cell) + pt_off
__global__ void __launch_bounds__(kBlock) k_cell_start(StepArgs a, BinWork w)
{
    const long long j = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (j >= w.n_cells_total) return;
    const int s = scene_of_cell(a.scenes, a.nscenes, j);
    const SceneDev &S = a.scenes[s];
    w.cell_start[j] = w.cell_count[j] - w.cell_count[S.cell_off] + (int)S.pt_off;
}

__global__ void __launch_bounds__(kBlock) k_scatter(StepArgs a, BinWork w)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= w.n_pts_total) return;
    const int bin = w.keys[i];
    if (bin < 0) return;
    const int s = scene_of_point(a.scenes, a.nscenes, i);
    const long long cell = a.scenes[s].cell_off + bin;
    const int slot = w.cell_start[cell] + atomicAdd(&w.cell_fill[cell], 1);
    w.perm[slot] = (int)i;
}

__global__ void __launch_bounds__(kBlock) k_sort_cells(StepArgs a, BinWork w)
{
    const long long j = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (j >= w.n_cells_total) return;
    const int s = scene_of_cell(a.scenes, a.nscenes, j);
    if (j - a.scenes[s].cell_off >= a.scenes[s].ncells) return;   // the per-scene end slot
    const int b = w.cell_start[j], e = w.cell_start[j + 1];
    for (int k = b + 1; k < e; k++) {
        const int v = w.perm[k];
        int m = k - 1;
        while (m >= b && w.perm[m] > v) { w.perm[m + 1] = w.perm[m]; m--; }
        w.perm[m + 1] = v;
    }
}

// Gathers the sorted points of every scene into dst; keys[] becomes the
// reference key of each sorted point.  Writes the new point counts.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_gather(StepArgs a, BinWork w, const T *src, T *dst, int *sorted_keys)
{
    const long long j = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (j >= w.n_pts_total) return;
    const int s = scene_of_point(a.scenes, a.nscenes, j);
    const SceneDev &S = a.scenes[s];
    const int kept = w.cell_start[S.cell_off + S.ncells] - (int)S.pt_off;
    if (j == S.pt_off) w.scene_count[s] = kept;
    if (j - S.pt_off >= kept) return;
    const int from = w.perm[j];
#pragma unroll
    for (int line = 0; line < kNumPtsLines; line++) dst[line * a.pitch + j] = src[line * a.pitch + from];
    const int bin = w.keys[from];
    const int col = bin / S.c.gridY, row = bin % S.c.gridY;
    sorted_keys[j] = (col - S.c.halo + S.c.xoff) * S.c.gridY + row;
}

// Points per tile (TX x TY cells) from cell_start; one thread per tile.
__global__ void __launch_bounds__(kBlock) k_tile_count(StepArgs a, const int *tile_off, int n_tiles_all, int *tile_n,
                                                       int *tile_flag)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_tiles_all) return;
    int lo = 0, hi = a.nscenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const SceneDev &S = a.scenes[lo];
    const int local = t - tile_off[lo];
    const int txi = local / S.n_tiles_y, tyi = local % S.n_tiles_y;
    const int col0 = txi * a.tx, row0 = tyi * a.ty;
    int n = 0;
    for (int c = 0; c < a.tx; c++) {
        const int col = col0 + c;
        if (col >= S.c.ncols) break;
        const long long base = S.cell_off + (long long)col * S.c.gridY;
        n += a.cell_start[base + min(row0 + a.ty, S.c.gridY)] - a.cell_start[base + row0];
    }
    tile_n[t] = n;
    tile_flag[t] = n > 0 ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_tile_emit(StepArgs a, const int *tile_off, int n_tiles_all,
                                                      const int *tile_n, const int *tile_pos, TileItem *tiles,
                                                      int *n_tiles)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t == n_tiles_all - 1) *n_tiles = tile_pos[t] + (tile_n[t] > 0 ? 1 : 0);
    if (t >= n_tiles_all || tile_n[t] == 0) return;
    int lo = 0, hi = a.nscenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const SceneDev &S = a.scenes[lo];
    const int local = t - tile_off[lo];
    const int col0 = (local / S.n_tiles_y) * a.tx, row0 = (local % S.n_tiles_y) * a.ty;
    const int rend = min(row0 + a.ty, S.c.gridY);
    // bounding box of the cells that hold points; cell_start is nondecreasing,
    // so the first and last occupied rows of a column are found by bisection
    int c_lo = a.tx, c_hi = -1, r_lo = a.ty, r_hi = -1;
    for (int k = 0; k < a.tx; k++) {
        const int col = col0 + k;
        if (col >= S.c.ncols) break;
        const int *cs = a.cell_start + S.cell_off + (long long)col * S.c.gridY;
        const int s0 = cs[row0], s1 = cs[rend];
        if (s1 == s0) continue;
        c_lo = min(c_lo, k);
        c_hi = k;
        int l = row0, h = rend - 1;           // first row r with cs[r + 1] > s0
        while (l < h) {
            const int m = (l + h) >> 1;
            if (cs[m + 1] > s0) h = m;
            else l = m + 1;
        }
        r_lo = min(r_lo, l - row0);
        l = row0;
        h = rend - 1;                         // last row r with cs[r] < s1
        while (l < h) {
            const int m = (l + h + 1) >> 1;
            if (cs[m] < s1) l = m;
            else h = m - 1;
        }
        r_hi = max(r_hi, l - row0);
    }
    TileItem it;
    it.scene = lo;
    it.col0 = col0 + c_lo;
    it.row0 = row0 + r_lo;
    it.dims = (c_hi - c_lo + 1) | ((r_hi - r_lo + 1) << 16);
    tiles[tile_pos[t]] = it;
}

inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// Exclusive scan of n ints (in may equal out), multi-level.
hipError_t scan_exclusive(const int *in, int *out, long long n, int *tmp, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const long long nb = (n + 1023) / 1024;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(1024), 0, s, in, out, nullptr, n);
        return hipGetLastError();
    }
    int *sums = tmp;
    hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(1024), 0, s, in, out, sums, n);
    hipError_t e = scan_exclusive(sums, sums, nb, tmp + nb, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(1024), 0, s, out, sums, n);
    return hipGetLastError();
}

hipError_t launch_sort(const StepArgs &a, const SceneDev *host_scenes, int nscenes, BinWork &w, int elem_bytes,
                       void *pts_src, void *pts_dst, hipStream_t s)
{
    (void)host_scenes;
    (void)nscenes;
    hipError_t e;
    const size_t cbytes = (size_t)w.n_cells_total * sizeof(int);
    if ((e = launch_fill_zero(w.cell_count, cbytes, s)) != hipSuccess) return e;
    if ((e = launch_fill_zero(w.cell_fill, cbytes, s)) != hipSuccess) return e;
    const unsigned gp = nblk(w.n_pts_total, kBlock), gc = nblk(w.n_cells_total, kBlock);
    if (elem_bytes == 8) hipLaunchKernelGGL(k_keys<8>, dim3(gp), dim3(kBlock), 0, s, a, w, pts_src);
    else hipLaunchKernelGGL(k_keys<4>, dim3(gp), dim3(kBlock), 0, s, a, w, pts_src);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = scan_exclusive(w.cell_count, w.cell_count, w.n_cells_total, w.scan_tmp, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_cell_start, dim3(gc), dim3(kBlock), 0, s, a, w);
    hipLaunchKernelGGL(k_scatter, dim3(gp), dim3(kBlock), 0, s, a, w);
    hipLaunchKernelGGL(k_sort_cells, dim3(gc), dim3(kBlock), 0, s, a, w);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(k_gather<double>, dim3(gp), dim3(kBlock), 0, s, a, w, (const double *)pts_src,
                           (double *)pts_dst, w.sorted_keys);
    else
        hipLaunchKernelGGL(k_gather<unsigned>, dim3(gp), dim3(kBlock), 0, s, a, w, (const unsigned *)pts_src,
                           (unsigned *)pts_dst, w.sorted_keys);
    return hipGetLastError();
}

hipError_t launch_build_tiles(const StepArgs &a, const SceneDev *host_scenes, int nscenes, const BinWork &w,
                              TileItem *tiles, int *n_tiles, long long max_tiles, hipStream_t s)
{
    (void)w;
    // tile_off, tile_n, tile_flag and tile_pos live in the tail of `tiles`
    // storage: the caller sized it as max_tiles TileItems followed by
    // 4*(max_tiles+nscenes+1) + max_tiles/1000 + 64 ints.
    int *ibase = (int *)(tiles + max_tiles);
    int *tile_off = ibase;
    int *tile_n = tile_off + nscenes + 1;
    int *tile_flag = tile_n + max_tiles;
    int *tile_pos = tile_flag + max_tiles;
    int *scan_tmp = tile_pos + max_tiles;
    long long total = 0;
    std::vector<int> off(nscenes + 1);
    for (int i = 0; i < nscenes; i++) {
        off[i] = (int)total;
        total += (long long)host_scenes[i].n_tiles_x * host_scenes[i].n_tiles_y;
    }
    off[nscenes] = (int)total;
    if (total > max_tiles || total == 0) return hipErrorInvalidValue;
    hipError_t e = hipMemcpyAsync(tile_off, off.data(), sizeof(int) * (nscenes + 1), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    const unsigned g = nblk(total, kBlock);
    hipLaunchKernelGGL(k_tile_count, dim3(g), dim3(kBlock), 0, s, a, tile_off, (int)total, tile_n, tile_flag);
    // positions of the non-empty tiles: exclusive scan of the 0/1 flags
    if ((e = scan_exclusive(tile_flag, tile_pos, total, scan_tmp, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_tile_emit, dim3(g), dim3(kBlock), 0, s, a, tile_off, (int)total, tile_n, tile_pos, tiles,
                       n_tiles);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the host copy of off[] must stay alive until the copy ran
    return hipStreamSynchronize(s);
}

}  // namespace icempm
