// Thread-local status and error text for the C-ABI.
#include <string>

#include "qotential_hip.h"

namespace qol {
thread_local int64_t g_status = 0;
thread_local std::string g_errmsg;
void set_status(int64_t s) { g_status = s; }
void set_error(const std::string& s) { g_errmsg = s; }
const char* error_text() { return g_errmsg.c_str(); }
}  // namespace qol

extern "C" int64_t qol_last_status(void) { return qol::g_status; }
