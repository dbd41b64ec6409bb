// flysim_amd: drop-in replacement of the reference's flysim.out command line.
#include <cstdio>

#include "../../../include/flysim.h"

int main(int argc, char **argv) {
  const int rc = fs_cli_main(argc, (const char *const *)argv);
  if (rc != FS_OK) {
    std::fprintf(stderr, "flysim_amd: error %d: %s\n", rc, fs_last_error());
    return 1;
  }
  return 0;
}
