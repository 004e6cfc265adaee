// `meteor`: drop-in replacement of the reference executable
// (source/makefile:6-9 builds it from meteor-fork.cpp; main() at :42).
// Like the reference it takes no arguments and works in the current
// directory.  MSC_PRECISION=64 selects the all-FP64 path (default 32).
#include <cstdio>
#include <cstdlib>

#include "../../include/msc.h"

int main() {
  int precision = 32;
  if (const char* s = std::getenv("MSC_PRECISION")) precision = std::atoi(s);
  int rc = msc_run_dir(".", precision);
  if (rc != MSC_OK) {
    std::fprintf(stderr, "meteor: %s\n", msc_last_error());
    return 1;
  }
  return 0;
}
