#include "host_common.hpp"

namespace fs {
static thread_local std::string g_last_error;
int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
}  // namespace fs

extern "C" const char *fs_last_error(void) { return fs::g_last_error.c_str(); }
