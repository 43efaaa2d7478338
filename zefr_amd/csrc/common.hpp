// common.hpp -- error handling shared by the host classes, the C ABI and the kernels.
//
// The reference turns errors into std::runtime_error("file:line:func: msg") through
// ThrowException (include/macros.hpp:52-54) and checks cudaGetLastError after each launch
// (include/macros.hpp:35-44).  Here host code throws zfr::Error; every C entry point
// catches it, stores the message in a thread-local buffer and returns a nonzero status.
#pragma once

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

namespace zfr {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_error(const std::string &msg);
const char *last_error();

// Record msg and return a failure status (for the kernel layer, which does not throw).
inline int fail(const std::string &msg) {
  set_error(msg);
  return 1;
}

inline int check_hip(hipError_t err, const char *what) {
  if (err == hipSuccess) return 0;
  set_error(std::string(what) + ": " + hipGetErrorString(err));
  return 2;
}

[[noreturn]] inline void raise(const std::string &msg) { throw Error(msg); }

inline void hip_or_throw(hipError_t err, const char *what) {
  if (err != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(err));
}

// Doubles of the operator-table region a residual kernel stages for elements with N
// points per direction (lE, Dss, Dsf: N^2 + 4 N, padded to 8).  A fused p-multigrid
// prolongation reuses that region for its [n_fine][N] table, so it is fused only where
// n_fine N fits.
inline int ops_region_doubles(int N) { return (N * N + 4 * N + 7) & ~7; }
inline bool pmg_prolong_fits(int N, int n_fine) { return n_fine >= N && n_fine <= 10 && n_fine * N <= ops_region_doubles(N); }

// The partner map packs (partner element, own slot, partner flux point) into 32 bits:
// ele << (B + 1) | slot << B | fpt, with B = 8 flux-point bits when an element has at most
// 256 flux points and 9 otherwise (hexes of order 6 to 8).
inline int nbr_fpt_bits(int nF) { return nF <= 256 ? 8 : 9; }

// Convert a kernel-layer status into an exception carrying its message.
inline void status_or_throw(int rc) {
  if (rc != 0) throw Error(last_error());
}

// Run fn at a C entry point: exceptions become a stored message and status 1.
template <typename Fn>
int guard(Fn &&fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception &ex) {
    set_error(ex.what());
    return 1;
  }
}

template <typename T, typename Fn>
T guard_val(T fail_value, Fn &&fn) {
  try {
    return fn();
  } catch (const std::exception &ex) {
    set_error(ex.what());
    return fail_value;
  }
}

}  // namespace zfr
