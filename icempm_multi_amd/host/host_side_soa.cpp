// host_side_soa.cpp -- HostSideSOA (simulation/hssoa/host_side_soa.cpp).
#include "host_side_soa.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <vector>

HostSideSOA::~HostSideSOA()
{
    if (host_buffer) (void)hipHostFree(host_buffer);
}

void HostSideSOA::Allocate(unsigned cap)
{
    if (host_buffer) { (void)hipHostFree(host_buffer); host_buffer = nullptr; }
    capacity = cap;
    const size_t bytes = sizeof(double) * (size_t)cap * icy::SimParams::nPtsArrays;
    if (hipHostMalloc((void **)&host_buffer, bytes, hipHostMallocDefault) != hipSuccess)
        throw std::runtime_error("allocating host buffer for points");
    size = 0;
    std::memset(host_buffer, 0, bytes);
}

int64_t HostSideSOA::utility(unsigned i) const
{
    int64_t v;
    std::memcpy(&v, host_buffer + i, 8);
    return v;
}

void HostSideSOA::setUtility(unsigned i, int64_t v) { std::memcpy(host_buffer + i, &v, 8); }

void HostSideSOA::setPartition(unsigned i, uint8_t pid)
{
    int64_t v = utility(i);
    v &= 0xffffffff00ffffffll;
    v |= (int64_t)pid << 24;
    setUtility(i, v);
}

int HostSideSOA::getCellIndex(unsigned i, double hinv, unsigned GridY) const
{
    const double x = host_buffer[(size_t)capacity * icy::SimParams::posx + i];
    const double y = host_buffer[(size_t)capacity * (icy::SimParams::posx + 1) + i];
    return (int)(x * hinv + 0.5) * (int)GridY + (int)(y * hinv + 0.5);
}

void HostSideSOA::RemoveDisabledAndSort(double hinv, unsigned GridY)
{
    std::vector<unsigned> idx;
    idx.reserve(size);
    for (unsigned i = 0; i < size; i++)
        if (!getDisabledStatus(i)) idx.push_back(i);
    std::vector<int> key(size);
    for (unsigned i : idx) key[i] = getCellIndex(i, hinv, GridY);
    std::stable_sort(idx.begin(), idx.end(), [&](unsigned a, unsigned b) { return key[a] < key[b]; });
    std::vector<double> tmp(idx.size());
    for (size_t line = 0; line < icy::SimParams::nPtsArrays; line++) {
        double *L = getPointerToLine((int)line);
        for (size_t k = 0; k < idx.size(); k++) tmp[k] = L[idx[k]];
        std::copy(tmp.begin(), tmp.end(), L);
    }
    size = (unsigned)idx.size();
}

void HostSideSOA::ApplyPermutation(const int32_t *perm, unsigned n)
{
    if (n > size) throw std::runtime_error("ApplyPermutation: more points than the buffer holds");
    for (unsigned k = 0; k < n; k++)
        if (perm[k] < 0 || (unsigned)perm[k] >= size) throw std::runtime_error("ApplyPermutation: index out of range");
    std::vector<double> tmp(n);
    for (size_t line = 0; line < icy::SimParams::nPtsArrays; line++) {
        double *L = getPointerToLine((int)line);
        for (unsigned k = 0; k < n; k++) tmp[k] = L[perm[k]];
        std::copy(tmp.begin(), tmp.end(), L);
    }
    size = n;
}

unsigned HostSideSOA::FindFirstPointAtGridXIndex(const int index_grid_x, const double hinv)
{
    const double *X = getPointerToPosX();
    unsigned lo = 0, hi = size;
    while (lo < hi) {
        const unsigned mid = lo + (hi - lo) / 2;
        if ((int)(X[mid] * hinv + 0.5) < index_grid_x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

std::pair<Vec2, Vec2> HostSideSOA::getBlockDimensions()
{
    const auto lo = getBlockMin(), hi = getBlockMax();
    return {Vec2(lo.first, lo.second), Vec2(hi.first, hi.second)};
}

void HostSideSOA::InitializeBlock()
{
    for (unsigned i = 0; i < size; i++) {
        getPointerToLine((int)icy::SimParams::idx_Jp_inv)[i] = 1;
        getPointerToLine((int)icy::SimParams::Fe00)[i] = 1.0;
        getPointerToLine((int)icy::SimParams::Fe00 + 3)[i] = 1.0;
    }
}

std::pair<double, double> HostSideSOA::getBlockMin() const
{
    const double *X = host_buffer + (size_t)capacity * icy::SimParams::posx;
    const double *Y = X + capacity;
    return {*std::min_element(X, X + size), *std::min_element(Y, Y + size)};
}

std::pair<double, double> HostSideSOA::getBlockMax() const
{
    const double *X = host_buffer + (size_t)capacity * icy::SimParams::posx;
    const double *Y = X + capacity;
    return {*std::max_element(X, X + size), *std::max_element(Y, Y + size)};
}

void HostSideSOA::offsetBlock(double dx, double dy)
{
    double *X = getPointerToPosX(), *Y = getPointerToPosY();
    for (unsigned i = 0; i < size; i++) { X[i] += dx; Y[i] += dy; }
}
