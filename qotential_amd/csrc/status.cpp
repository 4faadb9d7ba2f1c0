// Thread-local status and error text for the C-ABI.
#include <string>

#include "qotential_hip.h"

namespace qol {
thread_local int64_t g_status = 0;
thread_local std::string g_errmsg;
thread_local int64_t g_unconverged = 0;
thread_local double g_capped_step = 0.0;
void set_status(int64_t s) { g_status = s; }
void set_unconverged(int64_t n, double max_step) { g_unconverged = n; g_capped_step = max_step; }
void set_error(const std::string& s) { g_errmsg = s; }
const char* error_text() { return g_errmsg.c_str(); }
}  // namespace qol

extern "C" int64_t qol_last_status(void) { return qol::g_status; }
extern "C" int64_t qol_last_projection_unconverged(double* max_step) {
  if (max_step) *max_step = qol::g_capped_step;
  return qol::g_unconverged;
}
