// Host-side helpers shared by the library's translation units.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/msc.h"

namespace msc {

// Records `fmt` as the calling thread's last error and returns `code`.
int fail(int code, const char* fmt, ...);

// Constants of the C solver, meteor-fork.cpp:12-19.
constexpr double kPi = 3.141592653589793;
constexpr double kQE = 1.602177e-19;
constexpr double kE0 = 8.85419e-12;
constexpr double kU0 = 1.2566e-6;
constexpr double kME = 9.10939e-31;
constexpr double kC = 3.0000e8;

inline int pml_len(const msc_params* p) { return p->pmllen > 0 ? p->pmllen : 10; }

}  // namespace msc
