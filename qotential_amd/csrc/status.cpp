// Thread-local status for the drop-in entry points (whose reference signatures
// return nothing) and the error string behind qol_last_error().
#include <atomic>
#include <string>

#include "qotential_hip.h"

namespace qol {
thread_local int64_t g_status = 0;
void set_status(int64_t s) { g_status = s; }
}  // namespace qol

extern "C" int64_t qol_last_status(void) { return qol::g_status; }
