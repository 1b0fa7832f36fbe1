// pbamd_env.hh -- process-wide HIP context of the C++ class layer and the
// reference's error convention (cudaFrame::ErrorInfo prints and exits,
// cudaFrame.cu:20-32).
#pragma once
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include "../../../include/pbamd.h"

typedef float real;   // the reference is built with -DREAL=float (Makefile:25)

namespace pbamd {

inline int &current_device() {
  static int dev = 0;
  return dev;
}

inline pb_ctx *context() {
  static pb_ctx *ctx[64] = {nullptr};
  const int d = current_device();
  if (!ctx[d]) {
    int rc = pb_ctx_create(d, &ctx[d]);
    if (rc != PB_OK) {
      std::cerr << "pbamd: cannot open HIP device " << d << " (status " << rc << ")" << std::endl;
      std::exit(1);
    }
  }
  return ctx[d];
}

// maps a non-zero status to the reference's print-and-exit behaviour
inline void check(int rc, const std::string &what) {
  if (rc != PB_OK) {
    std::cerr << what << ": " << pb_ctx_last_error(context()) << " (status " << rc << ")"
              << std::endl;
    std::exit(1);
  }
}

}  // namespace pbamd
