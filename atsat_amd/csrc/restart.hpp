// The reference's binary restart format (cSaturnModel::save_state / load_state,
// planet/FileIO_Sat.cpp:513-660), shared by the host-array and the device entry points.
#pragma once

#include "atsat_gpu.h"

#include <functional>
#include <string>
#include <vector>

namespace atsat {

// one array of a restart file: a field of atsat_fields.h (field >= 0) or an extra (ATSAT_RX_*)
struct RestartItem { int field; int extra; };

// the arrays of a file of format 1, 2 or 3 in file order
std::vector<RestartItem> restart_items(int format);

// Writes the header and every array of the format; get(item, scratch, &data) points data at the
// array's im*jm*km doubles (its own storage or scratch, filled) and returns 0 or an error code.
int restart_write(const char* path, const atsat_grid& g, int iter, int format,
                  const std::function<int(const RestartItem&, double*, const double**)>& get, std::string& err);

// Checks the header against g, then for every array of the file's format: dest(item, scratch)
// returns where to read it (null: skip it), after(item, data) consumes it (0 or an error code).
int restart_read(const char* path, const atsat_grid& g, int* iter, int* format,
                 const std::function<double*(const RestartItem&, double*)>& dest,
                 const std::function<int(const RestartItem&, const double*)>& after, std::string& err);

// the library's last-error text (atsat_last_error(NULL)), set by calls without a context
void set_last_error(const std::string& msg);

}  // namespace atsat
