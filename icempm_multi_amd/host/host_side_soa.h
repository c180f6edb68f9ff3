// host_side_soa.h -- HostSideSOA (simulation/hssoa/host_side_soa.h:52-74): a
// pinned buffer of nPtsArrays lines of `capacity` doubles, and the
// utility_data accessors of ProxyPoint (proxypoint2d.cpp:90-176).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <utility>

#include "parameters_sim.h"

// A pair of doubles with the accessors callers use on the reference's
// Eigen::Vector2d results (x(), y(), [i]); Eigen is not a dependency here.
struct Vec2 {
    double v[2] = {0, 0};
    Vec2() = default;
    Vec2(double x, double y) : v{x, y} {}
    double x() const { return v[0]; }
    double y() const { return v[1]; }
    double &operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};

// One point of a HostSideSOA (proxypoint2d.h:8-38).  A reference (made by
// SOAIterator) reads and writes the point's slot in the 16 SoA lines; a value
// (any copy of a point) holds the 16 numbers itself.  Assigning to a reference
// writes through, so std algorithms that swap points through SOAIterator
// (std::sort, std::remove_if) permute the SoA lines.
struct ProxyPoint {
    constexpr static unsigned nArrays = icy::SimParams::nPtsArrays;
    bool isReference = false;
    unsigned pos = 0, pitch = 0;   // element index and capacity of each line
    double *soa = nullptr;         // nArrays lines of `pitch` doubles
    double data[nArrays] = {};     // the values of a non-reference point

    ProxyPoint() = default;
    ProxyPoint(const ProxyPoint &other) { *this = other; }
    ProxyPoint &operator=(const ProxyPoint &other)
    {
        for (unsigned i = 0; i < nArrays; i++) setValue(i, other.getValue(i));
        return *this;
    }

    double getValue(size_t idx) const { return isReference ? soa[pos + (size_t)pitch * idx] : data[idx]; }
    void setValue(size_t idx, double v)
    {
        if (isReference) soa[pos + (size_t)pitch * idx] = v;
        else data[idx] = v;
    }
    Vec2 getPos() const { return {getValue(icy::SimParams::posx), getValue(icy::SimParams::posx + 1)}; }
    Vec2 getVelocity() const { return {getValue(icy::SimParams::velx), getValue(icy::SimParams::velx + 1)}; }
    int64_t utility() const
    {
        const double d = getValue(icy::SimParams::idx_utility_data);
        int64_t v;
        std::memcpy(&v, &d, 8);
        return v;
    }
    void setUtility(int64_t v)
    {
        double d;
        std::memcpy(&d, &v, 8);
        setValue(icy::SimParams::idx_utility_data, d);
    }
    bool getCrushedStatus() const { return (utility() & ICEMPM_STATUS_CRUSHED) != 0; }
    bool getDisabledStatus() const { return (utility() & ICEMPM_STATUS_DISABLED) != 0; }
    bool getLiquidStatus() const { return (utility() & ICEMPM_STATUS_LIQUID) != 0; }
    void setLiquidStatus(bool val)
    {
        setUtility(val ? (utility() | ICEMPM_STATUS_LIQUID) : (utility() & ~(int64_t)ICEMPM_STATUS_LIQUID));
    }
    uint16_t getGrain() const { return (uint16_t)(utility() & 0xffff); }
    // the cell key of the sort: nearest node in x and y (proxypoint2d.cpp:164-169)
    int getCellIndex(double hinv, unsigned GridY) const
    {
        const Vec2 p = getPos();
        return (int)(p.x() * hinv + 0.5) * (int)GridY + (int)(p.y() * hinv + 0.5);
    }
    int getXIndex(double hinv) const { return (int)(getValue(icy::SimParams::posx) * hinv + 0.5); }
    void setPartition(uint8_t pid) { setUtility((utility() & 0xffffffff00ffffffll) | ((int64_t)pid << 24)); }
    uint8_t getPartition() const { return (uint8_t)((utility() >> 24) & 0xff); }
};

// Random-access iterator over the points of a HostSideSOA
// (host_side_soa.h:18-48); *it is a reference ProxyPoint.
class SOAIterator {
public:
    using iterator_category = std::random_access_iterator_tag;
    using value_type = ProxyPoint;
    using difference_type = int;
    using pointer = ProxyPoint *;
    using reference = ProxyPoint &;

    ProxyPoint m_point;

    SOAIterator() { m_point.isReference = true; }
    SOAIterator(unsigned pos, double *soa_data, unsigned pitch)
    {
        m_point.isReference = true;
        m_point.pos = pos;
        m_point.soa = soa_data;
        m_point.pitch = pitch;
    }
    // copies the position, never the point
    SOAIterator(const SOAIterator &o) : SOAIterator(o.m_point.pos, o.m_point.soa, o.m_point.pitch) {}
    SOAIterator &operator=(const SOAIterator &o)
    {
        m_point.pos = o.m_point.pos;
        m_point.soa = o.m_point.soa;
        m_point.pitch = o.m_point.pitch;
        return *this;
    }

    bool operator<(const SOAIterator &t) const { return m_point.pos < t.m_point.pos; }
    bool operator>(const SOAIterator &t) const { return m_point.pos > t.m_point.pos; }
    bool operator<=(const SOAIterator &t) const { return m_point.pos <= t.m_point.pos; }
    bool operator>=(const SOAIterator &t) const { return m_point.pos >= t.m_point.pos; }
    bool operator==(const SOAIterator &t) const { return m_point.pos == t.m_point.pos && m_point.soa == t.m_point.soa; }
    bool operator!=(const SOAIterator &t) const { return !(*this == t); }

    SOAIterator &operator+=(difference_type n) { m_point.pos += n; return *this; }
    SOAIterator &operator-=(difference_type n) { m_point.pos -= n; return *this; }
    SOAIterator &operator++() { ++m_point.pos; return *this; }
    SOAIterator &operator--() { --m_point.pos; return *this; }
    SOAIterator operator++(int) { SOAIterator r = *this; ++m_point.pos; return r; }
    SOAIterator operator--(int) { SOAIterator r = *this; --m_point.pos; return r; }
    SOAIterator operator+(difference_type n) const { SOAIterator r = *this; r.m_point.pos += n; return r; }
    SOAIterator operator-(difference_type n) const { SOAIterator r = *this; r.m_point.pos -= n; return r; }
    difference_type operator-(const SOAIterator &o) const { return (difference_type)m_point.pos - (difference_type)o.m_point.pos; }
    reference operator*() { return m_point; }
    pointer operator->() { return &m_point; }
    ProxyPoint operator[](difference_type n) const { return *(*this + n).operator->(); }
};

class HostSideSOA {
public:
    double *host_buffer = nullptr;
    unsigned capacity = 0;
    unsigned size = 0;

    HostSideSOA() = default;
    HostSideSOA(const HostSideSOA &) = delete;
    HostSideSOA &operator=(const HostSideSOA &) = delete;
    ~HostSideSOA();

    SOAIterator begin() { return SOAIterator(0, host_buffer, capacity); }
    SOAIterator end() { return SOAIterator(size, host_buffer, capacity); }

    void Allocate(unsigned capacity);
    void RemoveDisabledAndSort(double hinv, unsigned GridY);   // stable: ties keep their order
    // Point i of the result is point perm[i] of the current order, for i < n
    // (the permutation a device RemoveDisabledAndSort returns); size becomes n.
    void ApplyPermutation(const int32_t *perm, unsigned n);
    // first point whose x cell index is >= index_grid_x, in a sorted buffer
    // (host_side_soa.cpp:103-113)
    unsigned FindFirstPointAtGridXIndex(const int index_grid_x, const double hinv);
    void InitializeBlock();                                   // Fe = I, Jp_inv = 1
    std::pair<Vec2, Vec2> getBlockDimensions();               // (min x, min y), (max x, max y)
    std::pair<double, double> getBlockMin() const;
    std::pair<double, double> getBlockMax() const;
    void offsetBlock(double dx, double dy);
    void offsetBlock(Vec2 offset) { offsetBlock(offset.x(), offset.y()); }

    double *getPointerToPosX() { return host_buffer + (size_t)capacity * icy::SimParams::posx; }
    double *getPointerToPosY() { return host_buffer + (size_t)capacity * (icy::SimParams::posx + 1); }
    double *getPointerToLine(int idxLine) { return host_buffer + (size_t)capacity * idxLine; }

    int64_t utility(unsigned i) const;
    void setUtility(unsigned i, int64_t v);
    bool getDisabledStatus(unsigned i) const { return (utility(i) & ICEMPM_STATUS_DISABLED) != 0; }
    bool getCrushedStatus(unsigned i) const { return (utility(i) & ICEMPM_STATUS_CRUSHED) != 0; }
    bool getLiquidStatus(unsigned i) const { return (utility(i) & ICEMPM_STATUS_LIQUID) != 0; }
    uint16_t getGrain(unsigned i) const { return (uint16_t)(utility(i) & 0xffff); }
    uint8_t getPartition(unsigned i) const { return (uint8_t)((utility(i) >> 24) & 0xff); }
    void setPartition(unsigned i, uint8_t pid);
    int getCellIndex(unsigned i, double hinv, unsigned GridY) const;
};
