// parameters_sim.h -- icy::SimParams with the reference's methods
// (simulation/parameters_sim.h:19-105, parameters_sim.cpp), over the C-ABI's
// icempm_simparams POD (same members, same layout).
#pragma once

#include <string>

#include "../../include/icempm.h"

namespace icy {

struct SimParams : public icempm_simparams {
    constexpr static double pi = 3.14159265358979323846;
    constexpr static int dim = 2;
    constexpr static int nGridArrays = ICEMPM_N_GRID_ARRAYS;
    constexpr static size_t idx_utility_data = ICEMPM_IDX_UTILITY_DATA;
    constexpr static size_t idx_P = ICEMPM_IDX_P;
    constexpr static size_t idx_Q = ICEMPM_IDX_Q;
    constexpr static size_t idx_Jp_inv = ICEMPM_IDX_JP_INV;
    constexpr static size_t posx = ICEMPM_IDX_POSX;
    constexpr static size_t velx = ICEMPM_IDX_VELX;
    constexpr static size_t Fe00 = ICEMPM_IDX_FE00;
    constexpr static size_t Bp00 = ICEMPM_IDX_BP00;
    constexpr static size_t nPtsArrays = ICEMPM_N_PTS_ARRAYS;

    void Reset();                                   // parameters_sim.cpp:5-60
    std::string ParseFile(std::string fileName);    // parameters_sim.cpp:63-126, returns InputRawPoints
    void ComputeLame();                             // parameters_sim.cpp:128-133
    void ComputeCamClayParams2();                   // parameters_sim.cpp:155-165
    void ComputeHelperVariables();                  // parameters_sim.cpp:135-153
    void ComputeIntegerBlockCoords();               // parameters_sim.cpp:167-173
    double PointsPerCell() { return nPtsTotal / (Volume / (cellsize * cellsize)); }
    int AnimationFrameNumber() { return SimulationStep / UpdateEveryNthStep; }
    size_t IndenterArraySize() { return sizeof(double) * n_indenter_subdivisions * dim; }
    int CellIdx(float x) { return (int)(x * cellsize_inv + 0.5); }
    int PointCellIndex(float x, float y) { return CellIdx(x) * GridY + CellIdx(y); }
};

static_assert(sizeof(SimParams) == sizeof(icempm_simparams), "SimParams must stay a POD mirror");

}  // namespace icy
