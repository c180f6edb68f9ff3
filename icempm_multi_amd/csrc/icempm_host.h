// icempm_host.h -- host-side helpers shared by the runtime and the C-ABI.
#pragma once

#include <string>

#include "../../include/icempm.h"

namespace icempm {

void params_compute_lame(icempm_simparams *P);
void params_camclay(icempm_simparams *P);
void params_helper(icempm_simparams *P);
void params_reset(icempm_simparams *P);
int params_parse_json(icempm_simparams *P, const char *text, std::string &input_raw_points, std::string &err);
int prepare_raw_points(icempm_simparams *P, double *hssoa, unsigned capacity, int n, float volume,
                       std::string &err);

void set_last_error(const std::string &msg);

}  // namespace icempm
