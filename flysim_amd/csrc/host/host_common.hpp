// Error reporting shared by every C-ABI entry point: a thread-local message
// plus an int status; no exception ever crosses the ABI.
#pragma once
#include <exception>
#include <string>

#include "../../../include/flysim.h"

namespace fs {
int set_error(int code, const std::string &msg);
}

#define FS_TRY_BEGIN try {
#define FS_TRY_END                                                          \
  }                                                                         \
  catch (const std::bad_alloc &) {                                          \
    return fs::set_error(FS_ERR_CAPACITY, "host allocation failed");        \
  }                                                                         \
  catch (const std::exception &e) {                                         \
    return fs::set_error(FS_ERR_ARG, e.what());                             \
  }
