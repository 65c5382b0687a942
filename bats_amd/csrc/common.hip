// Error state and small host helpers of the BATS HIP library.
#include "common.hpp"

#include <string>

namespace bats {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(const std::string &msg) {
  set_error(msg);
  return 1;
}

int check_launch(const char *what) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(err));
  return 0;
}

}  // namespace bats

extern "C" const char *bats_last_error(void) { return bats::g_last_error.c_str(); }

extern "C" const char *bats_version(void) { return "bats_amd 0.1 (gfx950)"; }
