// host_side_soa.h -- HostSideSOA (simulation/hssoa/host_side_soa.h:52-74): a
// pinned buffer of nPtsArrays lines of `capacity` doubles, and the
// utility_data accessors of ProxyPoint (proxypoint2d.cpp:90-176).
#pragma once

#include <cstdint>
#include <utility>

#include "parameters_sim.h"

class HostSideSOA {
public:
    double *host_buffer = nullptr;
    unsigned capacity = 0;
    unsigned size = 0;

    HostSideSOA() = default;
    HostSideSOA(const HostSideSOA &) = delete;
    HostSideSOA &operator=(const HostSideSOA &) = delete;
    ~HostSideSOA();

    void Allocate(unsigned capacity);
    void RemoveDisabledAndSort(double hinv, unsigned GridY);   // stable: ties keep their order
    void InitializeBlock();                                   // Fe = I, Jp_inv = 1
    std::pair<double, double> getBlockMin() const;
    std::pair<double, double> getBlockMax() const;
    void offsetBlock(double dx, double dy);

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
