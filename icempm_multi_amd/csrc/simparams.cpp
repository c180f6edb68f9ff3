// simparams.cpp -- icy::SimParams host logic (simulation/parameters_sim.cpp:5-165)
// and the point preparation of SnapshotManager::LoadRawPoints
// (simulation/snapshotmanager.cpp:29-121), exported through the C-ABI.
//
// The reference parses its flat JSON with rapidjson; this image has none, so a
// minimal reader of flat objects (string / number / bool values) is used.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/icempm.h"
#include "icempm_host.h"

namespace {

constexpr double kPi = 3.14159265358979323846;   // parameters_sim.h:22

struct JsonValue {
    bool is_string = false;
    std::string s;
    double num = 0;
};

// Parses {"key": value, ...} with string, number, true/false/null values.
bool parse_flat_json(const char *text, std::map<std::string, JsonValue> &out, std::string &err)
{
    const char *p = text;
    auto ws = [&]() { while (*p && std::isspace((unsigned char)*p)) p++; };
    auto str = [&](std::string &dst) -> bool {
        if (*p != '"') return false;
        p++;
        dst.clear();
        while (*p && *p != '"') {
            if (*p == '\\' && p[1]) {
                p++;
                const char c = *p;
                dst.push_back(c == 'n' ? '\n' : c == 't' ? '\t' : c);
            } else {
                dst.push_back(*p);
            }
            p++;
        }
        if (*p != '"') return false;
        p++;
        return true;
    };
    ws();
    if (*p != '{') { err = "configuration file is not JSON"; return false; }
    p++;
    ws();
    if (*p == '}') return true;
    while (*p) {
        ws();
        std::string key;
        if (!str(key)) { err = "expected a key"; return false; }
        ws();
        if (*p != ':') { err = "expected ':' after " + key; return false; }
        p++;
        ws();
        JsonValue v;
        if (*p == '"') {
            if (!str(v.s)) { err = "bad string for " + key; return false; }
            v.is_string = true;
        } else if (!std::strncmp(p, "true", 4)) {
            v.num = 1; p += 4;
        } else if (!std::strncmp(p, "false", 5)) {
            v.num = 0; p += 5;
        } else if (!std::strncmp(p, "null", 4)) {
            p += 4;
        } else {
            char *end = nullptr;
            v.num = std::strtod(p, &end);
            if (end == p) { err = "bad value for " + key; return false; }
            p = end;
        }
        out[key] = v;
        ws();
        if (*p == ',') { p++; continue; }
        if (*p == '}') return true;
        err = "expected ',' or '}'";
        return false;
    }
    err = "unterminated object";
    return false;
}

}  // namespace

namespace icempm {

void params_compute_lame(icempm_simparams *P)
{
    P->lambda = P->YoungsModulus * P->PoissonsRatio / ((1 + P->PoissonsRatio) * (1 - 2 * P->PoissonsRatio));
    P->mu = P->YoungsModulus / (2 * (1 + P->PoissonsRatio));
    P->kappa = P->mu * 2. / 3. + P->lambda;
}

void params_camclay(icempm_simparams *P)
{
    params_compute_lame(P);
    P->NACC_beta = P->IceTensileStrength / P->IceCompressiveStrength;
    const double beta = P->NACC_beta, q = P->IceShearStrength, p0 = P->IceCompressiveStrength;
    P->NACC_M = (2 * q * std::sqrt(1 + 2 * beta)) / (p0 * (1 + beta));
    P->NACC_Msq = P->NACC_M * P->NACC_M;
}

void params_helper(icempm_simparams *P)
{
    P->ParticleMass = P->ParticleVolume * P->Density;
    P->WaterParticleMass = P->ParticleVolume * P->WaterDensity;
    P->UpdateEveryNthStep = (int)(1.f / (200 * P->InitialTimeStep));
    P->cellsize = P->GridXDimension / P->GridXTotal;
    P->cellsize_inv = 1. / P->cellsize;
    P->Dp_inv = 4. / (P->cellsize * P->cellsize);
    P->IndRSq = P->IndDiameter * P->IndDiameter / 4.;
    P->dt_vol_Dpinv = P->InitialTimeStep * P->ParticleVolume * P->Dp_inv;
    P->dt_Gravity = P->InitialTimeStep * P->Gravity;
    P->vmax = 0.5 * P->cellsize / P->InitialTimeStep;
    P->vmax_squared = P->vmax * P->vmax;
    P->VectorCapacity_transfer =
        P->nPartitions > 0 ? (int)(P->nPtsTotal / P->nPartitions * P->PointsTransferBufferFraction) : 0;
    P->gbOffset = P->GridY * P->GridHaloSize;
}

void params_reset(icempm_simparams *P)
{
    std::memset(P, 0, sizeof(*P));
    P->n_indenter_subdivisions = 100;
    P->InitialTimeStep = 3.e-5;
    P->YoungsModulus = 5.e8;
    P->GridXTotal = 128;
    P->GridY = 55;
    P->ParticleViewSize = 2.5f;
    P->GridXDimension = 3.33;
    P->SimulationEndTime = 12;
    P->PoissonsRatio = 0.3;
    P->Gravity = 9.81;
    P->Density = 916;
    P->WaterDensity = 1000;
    P->IndDiameter = 0.324;
    P->IndVelocity = 0.2;
    P->IndDepth = 0.25;
    P->IceCompressiveStrength = 100e6;
    P->IceTensileStrength = 10e6;
    P->IceShearStrength = 6e6;
    P->IceTensileStrength2 = 5e6;
    P->DP_tan_phi = std::tan(30 * kPi / 180.);
    P->DP_threshold_p = 1e3;
    P->tpb_P2G = 256;
    P->tpb_Upd = 512;
    P->tpb_G2P = 128;
    P->SetupType = 0;
    P->GrainVariability = 0.50;
    P->GridHaloSize = 5;
    P->ExtraSpaceForIncomingPoints = 0.2;
    P->PointsTransferBufferFraction = 0.05;
    P->nPartitions = 3;
    P->RebalanceThresholdFreeSpaceRemaining = 0.15;
    P->RebalanceThresholdDisabledPercentage = 0.01;
    P->PointTransferFrequency = 2;
    params_compute_lame(P);
    params_camclay(P);
    params_helper(P);
}

int params_parse_json(icempm_simparams *P, const char *text, std::string &input_raw_points, std::string &err)
{
    std::map<std::string, JsonValue> doc;
    if (!parse_flat_json(text, doc, err)) return ICEMPM_ERR_ARGUMENT;
    auto num = [&](const char *k, double &dst) {
        auto it = doc.find(k);
        if (it != doc.end() && !it->second.is_string) dst = it->second.num;
    };
    auto inum = [&](const char *k, int &dst) {
        auto it = doc.find(k);
        if (it != doc.end() && !it->second.is_string) dst = (int)it->second.num;
    };
    // parameters_sim.cpp:79-109 (the 28 keys)
    inum("SetupType", P->SetupType);
    num("InitialTimeStep", P->InitialTimeStep);
    num("YoungsModulus", P->YoungsModulus);
    inum("GridX", P->GridXTotal);
    inum("GridY", P->GridY);
    num("GridXDimension", P->GridXDimension);
    num("ParticleViewSize", P->ParticleViewSize);
    num("SimulationEndTime", P->SimulationEndTime);
    num("PoissonsRatio", P->PoissonsRatio);
    num("Gravity", P->Gravity);
    num("Density", P->Density);
    num("IndDiameter", P->IndDiameter);
    num("IndVelocity", P->IndVelocity);
    num("IndDepth", P->IndDepth);
    num("IceCompressiveStrength", P->IceCompressiveStrength);
    num("IceTensileStrength", P->IceTensileStrength);
    num("IceTensileStrength2", P->IceTensileStrength2);
    num("IceShearStrength", P->IceShearStrength);
    {
        auto it = doc.find("DP_phi");
        if (it != doc.end() && !it->second.is_string) P->DP_tan_phi = std::tan(it->second.num * kPi / 180);
    }
    num("DP_threshold_p", P->DP_threshold_p);
    num("GrainVariability", P->GrainVariability);
    inum("tpb_P2G", P->tpb_P2G);
    inum("tpb_Upd", P->tpb_Upd);
    inum("tpb_G2P", P->tpb_G2P);
    inum("GridHaloSize", P->GridHaloSize);
    inum("nPartitions", P->nPartitions);
    num("ExtraSpaceForIncomingPoints", P->ExtraSpaceForIncomingPoints);
    params_camclay(P);
    params_helper(P);
    auto it = doc.find("InputRawPoints");
    if (it == doc.end() || !it->second.is_string) {
        err = "config parameter missing: InputRawPoints";
        return ICEMPM_ERR_ARGUMENT;
    }
    input_raw_points = it->second.s;
    return ICEMPM_OK;
}

// SnapshotManager::LoadRawPoints after the file read (snapshotmanager.cpp:57-110):
// block extent, offset by ((GridX*h - L)/2, 2h), RemoveDisabledAndSort,
// InitializeBlock (Fe = I, Jp_inv = 1), indenter start, particle volume.
int prepare_raw_points(icempm_simparams *P, double *hssoa, unsigned capacity, int n, float volume,
                       std::string &err)
{
    if (n <= 0 || (unsigned)n > capacity) { err = "prepare_raw_points: bad point count"; return ICEMPM_ERR_ARGUMENT; }
    P->nPtsTotal = n;
    P->Volume = (double)volume;
    double *X = hssoa + (size_t)capacity * ICEMPM_IDX_POSX;
    double *Y = hssoa + (size_t)capacity * (ICEMPM_IDX_POSX + 1);
    P->xmin = *std::min_element(X, X + n);
    P->xmax = *std::max_element(X, X + n);
    P->ymin = *std::min_element(Y, Y + n);
    P->ymax = *std::max_element(Y, Y + n);
    const double h = P->cellsize;
    const double box_x = P->GridXTotal * h;
    const double length = P->xmax - P->xmin;
    const double x_offset = (box_x - length) / 2;
    const double y_offset = 2 * h;
    for (int i = 0; i < n; i++) { X[i] += x_offset; Y[i] += y_offset; }

    // RemoveDisabledAndSort (host_side_soa.cpp:50-61), ties in previous order
    std::vector<int> idx;
    idx.reserve(n);
    for (int i = 0; i < n; i++) {
        long long ud;
        std::memcpy(&ud, hssoa + i, 8);
        if (!(ud & ICEMPM_STATUS_DISABLED)) idx.push_back(i);
    }
    std::vector<int> key(n);
    for (int i : idx)
        key[i] = (int)(X[i] * P->cellsize_inv + 0.5) * P->GridY + (int)(Y[i] * P->cellsize_inv + 0.5);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
    const int m = (int)idx.size();
    std::vector<double> tmp(m);
    for (int line = 0; line < ICEMPM_N_PTS_ARRAYS; line++) {
        double *L = hssoa + (size_t)capacity * line;
        for (int k = 0; k < m; k++) tmp[k] = L[idx[k]];
        std::copy(tmp.begin(), tmp.end(), L);
    }
    // InitializeBlock (host_side_soa.cpp:94-105)
    for (int k = 0; k < m; k++) {
        hssoa[(size_t)capacity * ICEMPM_IDX_JP_INV + k] = 1;
        hssoa[(size_t)capacity * ICEMPM_IDX_FE00 + k] = 1.0;
        hssoa[(size_t)capacity * (ICEMPM_IDX_FE00 + 3) + k] = 1.0;
    }
    // indenter start (snapshotmanager.cpp:85-105)
    const double block_left = x_offset;
    const double block_top = P->ymax + y_offset;
    const double r = P->IndDiameter / 2;
    const double ht = r - P->IndDepth;
    const double x_ind_offset = std::sqrt(r * r - ht * ht);
    P->indenter_x = std::floor((block_left - x_ind_offset) / h) * h;
    if (P->SetupType == 0) P->indenter_y = block_top + ht;
    else if (P->SetupType == 1) P->indenter_y = std::ceil(block_top / h) * h;
    else P->indenter_x = P->indenter_y = -1;
    P->indenter_x_initial = P->indenter_x;
    P->indenter_y_initial = P->indenter_y;
    P->ParticleVolume = P->Volume / n;
    params_helper(P);
    return m;
}

}  // namespace icempm
