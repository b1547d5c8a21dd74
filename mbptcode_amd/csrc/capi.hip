// Library-wide C entry points: error text and ABI version.
#include "common.h"

namespace mbpt {
namespace {
thread_local std::string g_error;
}
void set_error(const std::string &msg) { g_error = msg; }
}  // namespace mbpt

extern "C" {

const char *mbpt_last_error(void) { return mbpt::g_error.c_str(); }

int mbpt_abi_version(void) { return 1; }

}  // extern "C"
