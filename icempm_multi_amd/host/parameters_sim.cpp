// parameters_sim.cpp -- icy::SimParams methods through the C-ABI.
#include "parameters_sim.h"

#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace icy {

void SimParams::Reset() { icempm_params_reset(this); }

std::string SimParams::ParseFile(std::string fileName)
{
    std::ifstream f(fileName);
    if (!f) throw std::runtime_error("configuration file is not found");
    std::stringstream ss;
    ss << f.rdbuf();
    char raw[4096];
    if (icempm_params_parse_json(this, ss.str().c_str(), raw, sizeof raw) != ICEMPM_OK)
        throw std::runtime_error(std::string("configuration file: ") + icempm_last_error());
    return std::string(raw);
}

void SimParams::ComputeLame()
{
    lambda = YoungsModulus * PoissonsRatio / ((1 + PoissonsRatio) * (1 - 2 * PoissonsRatio));
    mu = YoungsModulus / (2 * (1 + PoissonsRatio));
    kappa = mu * 2. / 3. + lambda;
}

void SimParams::ComputeCamClayParams2()
{
    ComputeLame();
    NACC_beta = IceTensileStrength / IceCompressiveStrength;
    NACC_M = (2 * IceShearStrength * std::sqrt(1 + 2 * NACC_beta)) / (IceCompressiveStrength * (1 + NACC_beta));
    NACC_Msq = NACC_M * NACC_M;
}

void SimParams::ComputeHelperVariables()
{
    // the C-ABI's derive runs ComputeCamClayParams2 first; keep the Lame values
    icempm_params_derive(this);
}

void SimParams::ComputeIntegerBlockCoords()
{
    nxmin = (int)std::floor(xmin / cellsize);
    nxmax = (int)std::ceil(xmax / cellsize);
    nymin = (int)std::floor(ymin / cellsize);
    nymax = (int)std::ceil(ymax / cellsize);
}

}  // namespace icy
