// evp1d_cache.hpp - cache hints for loads and stores that a launch makes once.
//
// MI355X keeps a 256 MiB last-level cache (Infinity Cache) in front of HBM.
// A kernel that streams through more data than that per launch evicts what
// the next launch reads again.  Marking the streams that a launch reads or
// writes once as non-temporal (global_load ... nt, global_store ... nt) is
// meant to let the cache keep the rest; the gain appears only in the size
// window that this explanation predicts (hints_pay), and no counter on the
// chip separates cache hits from DRAM reads.  Hints only: every load and
// store returns and writes the same values either way.
#pragma once

#include <hip/hip_runtime.h>

constexpr double kLastLevelCache = 256.0 * 1024 * 1024;

typedef double NtD2 __attribute__((ext_vector_type(2)));

template <bool NT>
__host__ __device__ __forceinline__ void out2(double2 *p, double2 v) {
  if (NT) {
    const NtD2 x = {v.x, v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<NtD2 *>(p));
  } else {
    *p = v;
  }
}
template <bool NT>
__host__ __device__ __forceinline__ void out1(double *p, double v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
template <bool NT>
__host__ __device__ __forceinline__ double2 in2(const double2 *p) {
  if (NT) {
    const NtD2 x = __builtin_nontemporal_load(reinterpret_cast<const NtD2 *>(p));
    return make_double2(x.x, x.y);
  }
  return *p;
}
template <bool NT>
__host__ __device__ __forceinline__ double in1(const double *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}

// The hints pay when all the bytes a launch touches exceed the cache but the
// part that every launch reads again about fits it (at most 1.3 times its
// size): below, everything fits anyway and the hints only push data out;
// above, the reread part is evicted either way.  Measured limits: the pair
// form gains 8 % at 221 and 273 MB reread and loses 4 % at 409 MB; the 2D
// fused form gains 14 % at about 328 MB and gains nothing at about 406 MB.
inline bool hints_pay(double reread_bytes, double all_bytes) {
  return all_bytes > kLastLevelCache && reread_bytes <= 1.3 * kLastLevelCache;
}
