// control.hpp — MEINV control file (Src/RINPUT.c:361-563 RDINPUT, keyword
// table :400-458, conversions of Get_Keys :48-357) -> engine configuration
// and the run settings (paths, solution box, chunking, caching).
#pragma once

#include <string>
#include <vector>

#include "../../include/meinv.h"

namespace mehost {

struct Control {
    me_config cfg;                       // inversion settings, internal units
    std::string data_path, wavelength_path, cache_path, result_path, fit_path;
    int sol_box[2][2] = {{-1, -1}, {-1, -1}};  // [x|y][first|last], -1: whole axis
    int nprofiles = 20;                  // pixels per chunk (cache granularity)
    bool cache_prof = false;             // read the whole box at once
    bool cache_inv = false;              // resumable progress file
    int verbose = 1;
    std::vector<std::string> lines;      // `lines` entries (only the first is used)
    std::vector<std::string> warnings;
};

// Parse a control file.  Returns false with a message for what the reference
// treats as fatal (missing required keywords, malformed numbers, lambda
// accept/reject below 1.5).
bool parse_control(const std::string &path, Control &out, std::string &err);
bool parse_control_text(const std::string &text, Control &out, std::string &err);

}  // namespace mehost
