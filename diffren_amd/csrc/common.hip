// Error reporting and version for the diffren_amd C ABI.
#include <string>

#include "dr_common.h"
#include "../../include/diffren_amd.h"

namespace dr {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
const char* get_error() { return g_last_error.c_str(); }
}  // namespace dr

extern "C" const char* dr_last_error(void) { return dr::get_error(); }
extern "C" const char* dr_version(void) { return "diffren_amd 0.1.0 gfx950"; }
