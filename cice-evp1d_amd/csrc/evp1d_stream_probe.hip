// HBM streaming probe: the rate at which this GPU moves bytes when every
// load is independent and coalesced, at a chosen read/write mix.  It is the
// ceiling the stencil kernels are compared against (their loads are gathers
// through neighbour lists); it computes nothing of the model.
//
// Each of n points reads nr double2 streams (16 bytes each, stream j at
// src + j * n), sums them and writes the sum to nw double2 streams.  With
// nw = 0 the store happens only for a value the inputs never hold, so the
// loads stay but nothing is written.  A thread handles `per` points, one
// grid's width apart (every access stays coalesced), with all their loads
// issued before the first sum: per = 4 keeps four times as many bytes in
// flight per thread.  With hints = 1 every load and store is non-temporal
// (the cache hints of evp1d_cache.hpp), the ceiling for kernels that use them.
//
//   int evp1d_probe_stream(int32_t nr, int32_t nw, int32_t per, int64_t n,
//                          int32_t reps, double *us_per_launch);
//   int evp1d_probe_stream_hints(int32_t nr, int32_t nw, int32_t per, int32_t hints,
//                                int64_t n, int32_t reps, double *us_per_launch);
//
// allocates (nr + nw) * n * 16 bytes, fills the inputs, runs two untimed
// launches and then reps timed ones (HIP events on its own stream), frees
// everything and returns 0, or -1 for an unsupported mix, -2 for a HIP
// error.  Supported: nr in 1..6 with nw 0 or 1, and nr = nw = 2; per 1 or 4
// with n a multiple of 256 * per; hints 0 or 1.  evp1d_probe_stream is
// hints = 0.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "evp1d_cache.hpp"

namespace {

// grid: n / (256 * PER) workgroups of 256 threads (n a multiple of 256 * PER)
template <int NR, int NW, int PER, bool NT>
__global__ __launch_bounds__(256) void k_stream(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                long n) {
  const long w = (long)gridDim.x * 256;
  const long i0 = (long)blockIdx.x * 256 + threadIdx.x;
  double2 v[PER][NR];
#pragma unroll
  for (int p = 0; p < PER; ++p)
#pragma unroll
    for (int j = 0; j < NR; ++j) v[p][j] = in2<NT>(src + ((long)j * n + i0 + p * w));
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    double2 s = make_double2(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      s.x += v[p][j].x;
      s.y += v[p][j].y;
    }
    const long i = i0 + p * w;
    if (NW == 0) {
      if (s.x == -1.0) out2<NT>(dst + i, s);   // inputs are >= 0: never taken
    } else {
#pragma unroll
      for (int j = 0; j < NW; ++j) out2<NT>(dst + ((long)j * n + i), make_double2(s.x + j, s.y));
    }
  }
}

__global__ void k_fill(double2 *p, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = make_double2((double)(i & 1023), 1.0);
}

using Fn = void (*)(const double2 *, double2 *, long);

template <int PER, bool NT>
Fn pick_per(int nr, int nw) {
  if (nw == 0) {
    static const Fn f[6] = {k_stream<1, 0, PER, NT>, k_stream<2, 0, PER, NT>, k_stream<3, 0, PER, NT>,
                            k_stream<4, 0, PER, NT>, k_stream<5, 0, PER, NT>, k_stream<6, 0, PER, NT>};
    return nr >= 1 && nr <= 6 ? f[nr - 1] : nullptr;
  }
  if (nw == 1) {
    static const Fn f[6] = {k_stream<1, 1, PER, NT>, k_stream<2, 1, PER, NT>, k_stream<3, 1, PER, NT>,
                            k_stream<4, 1, PER, NT>, k_stream<5, 1, PER, NT>, k_stream<6, 1, PER, NT>};
    return nr >= 1 && nr <= 6 ? f[nr - 1] : nullptr;
  }
  return nr == 2 && nw == 2 ? k_stream<2, 2, PER, NT> : nullptr;
}

Fn pick(int nr, int nw, int per, int hints) {
  if (hints != 0 && hints != 1) return nullptr;
  if (per == 1) return hints ? pick_per<1, true>(nr, nw) : pick_per<1, false>(nr, nw);
  if (per == 4) return hints ? pick_per<4, true>(nr, nw) : pick_per<4, false>(nr, nw);
  return nullptr;
}

}  // namespace

extern "C" int evp1d_probe_stream_hints(int32_t nr, int32_t nw, int32_t per, int32_t hints, int64_t n,
                                        int32_t reps, double *us_per_launch) {
  const Fn fn = pick(nr, nw, per, hints);
  if (!fn || n < 1 || n % (256L * per) != 0 || reps < 1 || !us_per_launch) return -1;
  double2 *src = nullptr, *dst = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -2;
  const unsigned grid = (unsigned)(n / (256L * per));
  float ms = 0.0f;
  if (hipMalloc((void **)&src, (size_t)nr * n * sizeof(double2)) != hipSuccess) goto done;
  if (hipMalloc((void **)&dst, (size_t)(nw > 0 ? nw : 1) * n * sizeof(double2)) != hipSuccess) goto done;
  if (hipStreamCreate(&st) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)(((long)nr * n + 255) / 256)), dim3(256), 0, st, src, (long)nr * n);
  for (int k = 0; k < 2; ++k) hipLaunchKernelGGL(fn, dim3(grid), dim3(256), 0, st, src, dst, (long)n);
  if (hipGetLastError() != hipSuccess) goto done;
  if (hipEventRecord(e0, st) != hipSuccess) goto done;
  for (int k = 0; k < reps; ++k) hipLaunchKernelGGL(fn, dim3(grid), dim3(256), 0, st, src, dst, (long)n);
  if (hipGetLastError() != hipSuccess) goto done;
  if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
  *us_per_launch = 1e3 * (double)ms / reps;
  rc = 0;
done:
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return rc;
}

extern "C" int evp1d_probe_stream(int32_t nr, int32_t nw, int32_t per, int64_t n, int32_t reps,
                                  double *us_per_launch) {
  return evp1d_probe_stream_hints(nr, nw, per, 0, n, reps, us_per_launch);
}
