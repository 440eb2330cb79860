// evp1d_device.hip - gfx950 kernels and the C ABI of libevp1d.so.
//
// Device layout (one handle = one 1D point list, see include/evp1d.h)
//
// Slot order.  The caller's 1D order (row-major from calc_2d_indices_init)
// is permuted on the device into four classes, each kept in caller order:
//     [ stress only | stress and stepu | stepu only | neither | ghosts ]
//       skipT=F       skipT=F            skipT=T      both T     na..navel-1
//       skipU=T       skipU=F            skipU=F                 (identity)
// so that stress runs over the contiguous slot range [0, t_end) and stepu
// over [u_begin, u_end) with no skip test and no partly used cache lines
// (with skips scattered through the caller's order, HBM moves every line
// that holds one active point).  Neighbour lists are stored in slot space;
// set/get of fields go through the permutation, so callers only ever see
// their own order.  The permutation is rebuilt when skip flags change.
//
//   vel   double2[navel]      (uvel, vvel): one 16-byte load per stress gather
//   sig   double[12][ldsig]   stressp_1..4, stressm_1..4, stress12_1..4, SoA
//   s15, s27, s36, s48 double2[navel]
//                             str1/str5, str2/str7, str3/str6, str4/str8: the
//                             pairs stepu gathers from one neighbour each
//                             (own, nw, sse, sw), so every gather is 16 bytes
//   idx   int32[6][na]        ee, ne, se, nw, sw, sse, slot space, 0-based
//   everything else           double[na] SoA, streamed once per kernel
//
// Kernels: one thread per slot.  stress and stepu are separate phases because
// stepu at a point gathers the str of three neighbours written by stress in
// the same subcycle and the next stress gathers the velocities of three
// neighbours written by stepu.  The subcycle loop stays on the device as a
// captured hipGraph (default) or as one cooperative kernel with grid barriers.
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/evp1d.h"
#include "evp1d_adjoint.hpp"
#include "evp1d_tiles.hpp"
#include "evp1d_point.hpp"
#include "evp1d_cache.hpp"
#include "evp1d_xcd.hpp"

#ifndef EVP1D_FMA
#define EVP1D_FMA 0
#endif
#ifndef EVP1D_BLOCK
#define EVP1D_BLOCK 256
#endif

namespace cg = cooperative_groups;
using evp1d::Scalars;

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local int g_err_code = 0;
static thread_local char g_err_msg[512] = "";

static int scratch_setup(int dev, bool quiet);   // evp1d_loss_noalloc's scratch address
static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err_msg, sizeof(g_err_msg), fmt, ap);
  va_end(ap);
  g_err_code = code;
  return code;
}

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(EVP1D_EHIP, "%s failed: %s (%s:%d)", #expr,                  \
                  hipGetErrorString(e_), __FILE__, __LINE__);                  \
  } while (0)

#define RC_TRY(expr)              \
  do {                            \
    const int rc_ = (expr);       \
    if (rc_) return rc_;          \
  } while (0)

// ---------------------------------------------------------------------------
// device view of a handle (passed by value to every kernel)
// ---------------------------------------------------------------------------
struct DevView {
  int na, navel;
  long ldsig;
  int t_end;              // stress slots [0, t_end)
  int u_begin, u_end;     // stepu slots [u_begin, u_end)
  const int32_t *ee, *ne, *se, *nw, *sw, *sse;
  double2 *vel;
  double *sig;
  double2 *s15, *s27, *s36, *s48;
  // Split addressing (gradient trajectories): vel holds only the stepu slots
  // [u_begin, u_end); the other slots, whose values never change during a
  // run, are read from velc.  The str pairs may be split the same way over
  // the stress slots [0, t_end), with s15c..s48c for the rest (the gradient
  // sweep passes the handle's own arrays as both).  vel_out receives stepu's
  // result.  Unsplit kernels use vel for everything and vel_out == vel.
  double2 *vel_out;
  const double2 *velc, *s15c, *s27c, *s36c, *s48c;
  const double *dxT, *dyT, *strength, *hte, *htn, *htem1, *htnm1;
  const double *cw, *aiu, *uocn, *vocn, *waterx, *watery, *forcex, *forcey;
  const double *umassdti, *fm, *uarear, *uvel_init, *vvel_init, *tbu;
  double *strintx, *strinty, *cb;
  int nhalo;
  const int32_t *halo_dst, *halo_src;
  Scalars s;
};

template <bool SPLIT>
struct Ld {
  __device__ static __forceinline__ double2 vel(const DevView &d, int j) {
    if (!SPLIT) return d.vel[j];
    return (j >= d.u_begin && j < d.u_end) ? d.vel[j - d.u_begin] : d.velc[j];
  }
  __device__ static __forceinline__ double2 pair(const double2 *a, const double2 *c, const DevView &d, int j) {
    if (!SPLIT) return a[j];
    return j < d.t_end ? a[j] : c[j];
  }
  __device__ static __forceinline__ int vslot(const DevView &d, int j) { return SPLIT ? j - d.u_begin : j; }
};

// stress at slot iw from the view's velocities and the stresses at sp[c * ss],
// c = 0..11 (no stores)
template <bool SPLIT = false>
__device__ __forceinline__ void stress_eval_from(const DevView &d, int iw, const double *sp, long ss,
                                                 evp1d::StressOut &o) {
  using L = Ld<SPLIT>;
  const double2 cc = L::vel(d, iw);
  const double2 ve = L::vel(d, d.ee[iw]);
  const double2 vn = L::vel(d, d.ne[iw]);
  const double2 vs = L::vel(d, d.se[iw]);
  double sig[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) sig[c] = sp[c * ss];
  evp1d::stress_point(d.s, cc.x, cc.y, ve.x, ve.y, vs.x, vs.y, vn.x, vn.y,
                      d.dxT[iw], d.dyT[iw], d.strength[iw], d.hte[iw], d.htn[iw],
                      d.htem1[iw], d.htnm1[iw], sig, o);
}

// stress at slot iw from the view's velocities and stresses (no stores)
template <bool SPLIT = false>
__device__ __forceinline__ void stress_eval(const DevView &d, int iw, evp1d::StressOut &o) {
  stress_eval_from<SPLIT>(d, iw, d.sig + iw, d.ldsig, o);
}

template <bool SPLIT = false>
__device__ __forceinline__ void stress_at(const DevView &d, int iw) {
  evp1d::StressOut o;
  stress_eval<SPLIT>(d, iw, o);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    d.sig[c * d.ldsig + iw] = o.sp[c];
    d.sig[(4 + c) * d.ldsig + iw] = o.sm[c];
    d.sig[(8 + c) * d.ldsig + iw] = o.s12[c];
  }
  d.s15[iw] = make_double2(o.str[0], o.str[4]);
  d.s27[iw] = make_double2(o.str[1], o.str[6]);
  d.s36[iw] = make_double2(o.str[2], o.str[5]);
  d.s48[iw] = make_double2(o.str[3], o.str[7]);
}

template <bool SPLIT = false>
__device__ __forceinline__ void stepu_at(const DevView &d, int iw, bool write_diag) {
  using L = Ld<SPLIT>;
  const double2 v = L::vel(d, iw);
  const double2 a = L::pair(d.s15, d.s15c, d, iw);
  const double2 b = L::pair(d.s27, d.s27c, d, d.nw[iw]);
  const double2 c = L::pair(d.s36, d.s36c, d, d.sse[iw]);
  const double2 e = L::pair(d.s48, d.s48c, d, d.sw[iw]);
  evp1d::StepuOut o;
  evp1d::stepu_point(d.s, v.x, v.y, d.cw[iw], d.aiu[iw], d.uocn[iw], d.vocn[iw],
                     d.waterx[iw], d.watery[iw], d.forcex[iw], d.forcey[iw],
                     d.umassdti[iw], d.fm[iw], d.uarear[iw], d.uvel_init[iw],
                     d.vvel_init[iw], d.tbu[iw],
                     a.x, b.x, c.x, e.x, a.y, c.y, b.y, e.y, o);
  d.vel_out[L::vslot(d, iw)] = make_double2(o.uvel, o.vvel);
  // strintx/strinty/Cb are overwritten every subcycle and read by nothing
  // inside the loop; only the values of the final subcycle are observable.
  if (write_diag) {
    d.strintx[iw] = o.strintx;
    d.strinty[iw] = o.strinty;
    d.cb[iw] = o.Cb;
  }
}

__device__ __forceinline__ void halo_at(const DevView &d, int k) {
  const int dst = d.halo_dst[k];
  const int src = d.halo_src[k];
  d.vel[dst] = (src >= 0) ? d.vel[src] : make_double2(0.0, 0.0);
}

constexpr int kBlock = EVP1D_BLOCK;   // block size of the helper kernels and the persistent form

// tile_of (evp1d_xcd.hpp): the XCD-aware workgroup-to-tile mapping; for the
// unfused kernels a "tile" is a block of B slots, so each XCD streams one
// contiguous eighth of the slot range (the previous row for stress, the next
// row for stepu come through the same L2).

template <int B, bool X>
__global__ __launch_bounds__(B) void k_stress(DevView d) {
  const int iw = tile_of(blockIdx.x, gridDim.x, X) * B + threadIdx.x;
  if (iw < d.t_end) stress_at(d, iw);
}

template <int B, bool X>
__global__ __launch_bounds__(B) void k_stepu(DevView d, int write_diag) {
  const int iw = d.u_begin + tile_of(blockIdx.x, gridDim.x, X) * B + threadIdx.x;
  if (iw < d.u_end) stepu_at(d, iw, write_diag != 0);
}

// split-addressed forward kernels of the stored-trajectory gradient sweep
__global__ __launch_bounds__(256) void k_stress_split(DevView d) {
  const int iw = blockIdx.x * 256 + threadIdx.x;
  if (iw < d.t_end) stress_at<true>(d, iw);
}
__global__ __launch_bounds__(256) void k_stepu_split(DevView d, int write_diag) {
  const int iw = d.u_begin + blockIdx.x * 256 + threadIdx.x;
  if (iw < d.u_end) stepu_at<true>(d, iw, write_diag != 0);
}

// ---------------------------------------------------------------------------
// fused subcycle: stress and stepu of one tile of points per workgroup
// ---------------------------------------------------------------------------
// Tile t owns the stress slots [a0,a1) and [b0,b1) and the stepu slots
// [b0,b1) and [c0,c1).  Its LDS table holds str1..8 of [owned | halo |
// constant] entries: owned and halo entries are evaluated here from the
// input velocities and stresses (halo entries belong to other tiles and are
// not stored), constant entries (points whose str1..8 stress never writes)
// come from cstr.  Stepu then reads its own, nw, sse and sw entries from LDS.
// Velocities ping-pong between two buffers (in: d.vel, out: f.vel_out), so
// neighbour tiles never see a half-updated subcycle.  Stresses either do the
// same (in: d.sig, out: f.sig_out) or, with f.inplace, are updated in place
// (f.sig_out == d.sig): only their owner reads them, except for the "shared"
// points that other tiles evaluate as halo.  Those live during an in-place
// run in a double-buffered copy indexed by shared id: every evaluation reads
// hsig_in (the previous subcycle), their owner writes hsig_out, and
// k_scatter_shared puts the last copy back into d.sig at the end of the run.
#ifndef EVP1D_SH_STRIDE
#define EVP1D_SH_STRIDE 12
#endif
constexpr int kShStride = EVP1D_SH_STRIDE;   // doubles per shared point (12 stresses, point-major)
struct FusedView {
  const evp1d::TileDesc *tiles;
  const int32_t *hlist;     // halo slots of every tile
  const int32_t *hid;       // shared id of every halo entry (parallel to hlist)
  const uint2 *uloc;        // per stepu slot (slot - u_begin): own | nw << 16, sse | sw << 16
  const double2 *cstr;      // [4][nk]: s15, s27, s36, s48 of the constant entries
  const int *cnz;           // nonzero when some constant entry has a bit set (else all are +0.0)
  long nk;
  int inplace;              // stresses in place (shared points through hsig)
  long nsh;                 // shared points
  const double *hsig_in;    // [nsh][kShStride]
  double *hsig_out;         // [nsh][kShStride]
  double2 *vel_out;
  double *sig_out;          // == d.sig when inplace
  int cap;                  // entries per LDS table
};

// SPLIT: the gradient trajectory's forward sweep.  Velocities are split
// addressed (see Ld) and stepu writes f.vel_out[q - u_begin], one subcycle's
// entries of the trajectory.
template <int B, bool X, bool SPLIT = false>
__global__ __launch_bounds__(B) void k_fused(DevView d, FusedView f, int last) {
  extern __shared__ double2 lds[];
  const int cap = f.cap;
  const evp1d::TileDesc t = f.tiles[tile_of(blockIdx.x, gridDim.x, X)];
  const int nA = t.a1 - t.a0, nB = t.b1 - t.b0, nO = nA + nB, nS = nO + t.nh, nE = nS + t.nk;
  const bool cz = *f.cnz == 0;   // constant entries all +0.0: no loads
  for (int k = threadIdx.x; k < nE; k += B) {
    double2 p15, p27, p36, p48;
    if (k < nS) {
      // stresses from d.sig, shared points of an in-place run (owned ones
      // lead their range) from copy in; written to f.sig_out or copy out
      int g, id;
      if (k < nO) {
        const bool a = k < nA;
        const int r = a ? k : k - nA;
        g = (a ? t.a0 : t.b0) + r;
        id = f.inplace && r < (a ? t.sa : t.sb) ? (a ? t.ia : t.ib) + r : -1;
      } else {
        const int j = t.h0 + (k - nO);
        g = f.hlist[j];
        id = f.inplace ? f.hid[j] : -1;
      }
      const double *sp = id >= 0 ? f.hsig_in + (long)id * kShStride : d.sig + g;
      const long ss = id >= 0 ? 1 : d.ldsig;
      evp1d::StressOut o;
      stress_eval_from<SPLIT>(d, g, sp, ss, o);
      p15 = make_double2(o.str[0], o.str[4]);
      p27 = make_double2(o.str[1], o.str[6]);
      p36 = make_double2(o.str[2], o.str[5]);
      p48 = make_double2(o.str[3], o.str[7]);
      if (k < nO) {
        double *wp = id >= 0 ? f.hsig_out + (long)id * kShStride : f.sig_out + g;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          wp[c * ss] = o.sp[c];
          wp[(4 + c) * ss] = o.sm[c];
          wp[(8 + c) * ss] = o.s12[c];
        }
        if (last) {   // str1..8 are observable after the loop
          d.s15[g] = p15;
          d.s27[g] = p27;
          d.s36[g] = p36;
          d.s48[g] = p48;
        }
      }
    } else if (cz) {
      p15 = p27 = p36 = p48 = make_double2(0.0, 0.0);
    } else {
      const long j = t.k0 + (k - nS);
      p15 = f.cstr[j];
      p27 = f.cstr[f.nk + j];
      p36 = f.cstr[2 * f.nk + j];
      p48 = f.cstr[3 * f.nk + j];
    }
    lds[k] = p15;
    lds[cap + k] = p27;
    lds[2 * cap + k] = p36;
    lds[3 * cap + k] = p48;
  }
  __syncthreads();
  const int nU = nB + (t.c1 - t.c0);
  for (int k = threadIdx.x; k < nU; k += B) {
    const int q = k < nB ? t.b0 + k : t.c0 + (k - nB);
    const uint2 w = f.uloc[q - d.u_begin];
    const double2 a = lds[w.x & 0xffffu];
    const double2 b = lds[cap + (w.x >> 16)];
    const double2 c = lds[2 * cap + (w.y & 0xffffu)];
    const double2 e = lds[3 * cap + (w.y >> 16)];
    const double2 v = Ld<SPLIT>::vel(d, q);
    evp1d::StepuOut o;
    evp1d::stepu_point(d.s, v.x, v.y, d.cw[q], d.aiu[q], d.uocn[q], d.vocn[q], d.waterx[q], d.watery[q],
                       d.forcex[q], d.forcey[q], d.umassdti[q], d.fm[q], d.uarear[q], d.uvel_init[q],
                       d.vvel_init[q], d.tbu[q], a.x, b.x, c.x, e.x, a.y, c.y, b.y, e.y, o);
    f.vel_out[Ld<SPLIT>::vslot(d, q)] = make_double2(o.uvel, o.vvel);
    if (last) {
      d.strintx[q] = o.strintx;
      d.strinty[q] = o.strinty;
      d.cb[q] = o.Cb;
    }
  }
}

// ---------------------------------------------------------------------------
// two subcycles per workgroup ("pair" form)
// ---------------------------------------------------------------------------
// Level 1 runs stress on the tile's T1 entries (owned, halo, T1-only) from the
// input state d.vel / d.sig, then stepu on its U1 entries into an LDS
// velocity table; level 2 runs stress on the owned and halo entries from
// those velocities and the level-1 stresses, then stepu on the owned U
// points, and writes the state two subcycles on to vel_out / sig_out (two
// buffers each).  Every point computes exactly what the one-subcycle kernels
// compute, so results are bit-identical; neighbour tiles evaluate the border
// rings again instead of exchanging them, which halves the loads and stores
// of the state and the constant fields per subcycle.  Level-2 stress entry k
// is thread k (the plan guarantees at most B of them), so its level-1
// stresses stay in that thread's registers.  The 20 constant fields come
// from a copy packed in pairs at the start of the run (k_pack_const), the
// stresses live during the run in pairs too (k_pack_sig / k_unpack_sig): one
// 16-byte load per pair, since the kernel's time follows its load
// instructions.  Each level-1 stress entry's slot and ee, ne, se neighbour
// slots come from one per-tile int4 (PairPlan::tnbr), so a halo or T1-only
// entry reaches its velocities after two dependent loads, not three.
struct PairView {
  const evp1d::PairTile *tiles;
  const int32_t *hlist, *xlist, *elist;
  const uint2 *uloc;        // per U1 entry: own | nw << 16, sse | sw << 16 (str table entries)
  const uint2 *vloc;        // per T2 entry: cc | ee << 16, ne | se << 16 (velocity table entries)
  const double2 *cstr;      // [4][nk]: s15, s27, s36, s48 of the constant entries
  const int *cnz;
  long nk;
  double2 *vel_out;
  int cap;                  // entries of the str table (the velocity table follows it)
  const double2 *pc;        // [kPairConst][lp]: the constant fields in pairs (k_pack_const)
  long lp;
  const double2 *sg_in;     // [6][ls]: stresses in pairs (p1 p2, p3 p4, m1 m2, m3 m4, 12_1 12_2, 12_3 12_4)
  double2 *sg_out;
  long ls;
  const int4 *tnbr;         // per level-1 stress entry (PairTile::n0 + k): slot, ee, ne, se
  const int32_t *fix;       // fixed copy-list sources of level-2 ghost reads (PairPlan::fix)
  // gradient trajectory (k_pair<.., TRAJ>): the velocities of the stepu
  // slots after the launch's first subcycle (tr_mid) and after its second
  // (tr_end, null when the handle's own buffer is the only copy wanted),
  // indexed by slot - u_begin
  double2 *tr_mid, *tr_end;
};

// the stresses of the stress slots [0, n) into pairs, and back
__global__ void k_pack_sig(double2 *p, long ls, const double *sig, long ldsig, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) p[c * ls + i] = make_double2(sig[2 * c * ldsig + i], sig[(2 * c + 1) * ldsig + i]);
}
__global__ void k_unpack_sig(double *sig, long ldsig, const double2 *p, long ls, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double2 v = p[c * ls + i];
    sig[2 * c * ldsig + i] = v.x;
    sig[(2 * c + 1) * ldsig + i] = v.y;
  }
}
__global__ void k_pack_nbr(DevView d, int4 *nbr, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  nbr[i] = make_int4(d.ee[i], d.ne[i], d.se[i], 0);
}

// The constant fields in pairs, one double2 array each: (dxT, dyT), (hte,
// htn), (htem1, htnm1), (cw, aiu), (uocn, vocn), (waterx, watery), (forcex,
// forcey), (umassdti, fm), (uarear, tbu), (uvel_init, vvel_init).
constexpr int kPairConst = 10;
__global__ void k_pack_const(DevView d, double2 *pc, long lp) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.na) return;
  pc[i] = make_double2(d.dxT[i], d.dyT[i]);
  pc[lp + i] = make_double2(d.hte[i], d.htn[i]);
  pc[2 * lp + i] = make_double2(d.htem1[i], d.htnm1[i]);
  pc[3 * lp + i] = make_double2(d.cw[i], d.aiu[i]);
  pc[4 * lp + i] = make_double2(d.uocn[i], d.vocn[i]);
  pc[5 * lp + i] = make_double2(d.waterx[i], d.watery[i]);
  pc[6 * lp + i] = make_double2(d.forcex[i], d.forcey[i]);
  pc[7 * lp + i] = make_double2(d.umassdti[i], d.fm[i]);
  pc[8 * lp + i] = make_double2(d.uarear[i], d.tbu[i]);
  pc[9 * lp + i] = make_double2(d.uvel_init[i], d.vvel_init[i]);
}

// the buffers that alternate between launches: input velocities and
// stresses, output velocities and stresses
struct PairBufs {
  const double2 *vel;
  double2 *vel_out;
  const double2 *sg_in;
  double2 *sg_out;
};

// At most 128 VGPRs (4 waves per SIMD): workgroups of 256, 512 or 1024
// threads then fill 16 waves per CU.  TRAJ: the gradient's forward sweep,
// which also stores both subcycles' stepu-slot velocities (f.tr_mid,
// f.tr_end); the values computed are the same.
//
// Each barrier waits for LDS only, so global loads issued before it stay in
// flight across it: every thread loads the memory inputs of its first entry
// of the next phase (level-1 stepu, level-2 stress, level-2 stepu) before
// the barrier that ends the current phase, and that phase starts on values
// already in registers instead of one memory round trip after the barrier.
//
// NT: the cache hints (evp1d_cache.hpp).  The streams a launch reads or
// writes once (stresses in and out, the T points' constant fields,
// velocities and trajectory written) become non-temporal loads and stores,
// so that the cache can keep what every launch reads again (the U points'
// constant fields, velocities, tables).
template <int B, bool TRAJ, bool FIX, bool NT>
__device__ __forceinline__ void pair_tile(const DevView &d, const PairView &f, PairBufs bf, int last, int tix) {
  extern __shared__ double2 lds[];
  const int cap = f.cap;
  double2 *vt = lds + 4 * cap;
  const evp1d::PairTile pt = f.tiles[tix];
  const evp1d::TileDesc &t = pt.t;
  const int nA = t.a1 - t.a0, nB = t.b1 - t.b0, nO = nA + nB, nS = nO + t.nh, nT = nS + pt.nx, nE = nT + pt.nq;
  const int nUo = nB + (t.c1 - t.c0), nU = nUo + pt.ne;
  const bool cz = *f.cnz == 0;
  auto tslot = [&](int k) {
    return k < nA ? t.a0 + k : k < nO ? t.b0 + (k - nA) : k < nS ? f.hlist[t.h0 + (k - nO)] : f.xlist[pt.x0 + (k - nS)];
  };
  auto uslot = [&](int k) { return k < nB ? t.b0 + k : k < nUo ? t.c0 + (k - nB) : f.elist[pt.e0 + (k - nUo)]; };
  auto put = [&](int k, const evp1d::StressOut &o) {
    lds[k] = make_double2(o.str[0], o.str[4]);
    lds[cap + k] = make_double2(o.str[1], o.str[6]);
    lds[2 * cap + k] = make_double2(o.str[2], o.str[5]);
    lds[3 * cap + k] = make_double2(o.str[3], o.str[7]);
  };
  const double2 *pc = f.pc;
  const long lp = f.lp;
  // memory inputs of one stepu evaluation: its str table entries and the
  // seven packed constant pairs of slot q
  struct UIn {
    uint2 w;
    double2 c[7];
  };
  auto uload = [&](int k, int q) {
    UIn u;
    u.w = f.uloc[pt.u0 + k];
#pragma unroll
    for (int c = 0; c < 7; ++c) u.c[c] = pc[(3 + c) * lp + q];
    return u;
  };
  auto stepu = [&](const UIn &u, double2 v, evp1d::StepuOut &o) {
    const double2 a = lds[u.w.x & 0xffffu];
    const double2 b = lds[cap + (u.w.x >> 16)];
    const double2 c = lds[2 * cap + (u.w.y & 0xffffu)];
    const double2 e = lds[3 * cap + (u.w.y >> 16)];
    const double2 *k = u.c;
    evp1d::stepu_point(d.s, v.x, v.y, k[0].x, k[0].y, k[1].x, k[1].y, k[2].x, k[2].y, k[3].x, k[3].y, k[4].x,
                       k[4].y, k[5].x, k[6].x, k[6].y, k[5].y, a.x, b.x, c.x, e.x, a.y, c.y, b.y, e.y, o);
  };
  const int me = threadIdx.x;
  double sg[12];   // level-1 stresses of entry `me` when me < nS
  // entry `me`'s metric pairs and strength, loaded at level 1 and used again
  // at level 2 (the same slot)
  double2 m2[3];
  double st2 = 0.0;
  // level 1, stress (or constants) of entry k
  auto level1 = [&](int k) {
    if (k < nT) {
      const int4 n = f.tnbr[pt.n0 + k];   // slot and ee, ne, se of the entry
      const int g = k < nO ? tslot(k) : n.x;
      double sig[12];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const double2 v = in2<NT>(bf.sg_in + (c * f.ls + g));
        sig[2 * c] = v.x;
        sig[2 * c + 1] = v.y;
      }
      const double2 a0 = in2<NT>(pc + g), a1 = in2<NT>(pc + (lp + g)), a2 = in2<NT>(pc + (2 * lp + g));
      const double st = in1<NT>(d.strength + g);
      evp1d::StressOut o;
      const double2 cc = bf.vel[g], ve = bf.vel[n.y], vn = bf.vel[n.z], vs = bf.vel[n.w];
      evp1d::stress_point(d.s, cc.x, cc.y, ve.x, ve.y, vs.x, vs.y, vn.x, vn.y, a0.x, a0.y, st, a1.x, a1.y, a2.x,
                          a2.y, sig, o);
      if (k == me) {
        m2[0] = a0;
        m2[1] = a1;
        m2[2] = a2;
        st2 = st;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          sg[c] = o.sp[c];
          sg[4 + c] = o.sm[c];
          sg[8 + c] = o.s12[c];
        }
      }
      put(k, o);
    } else {
      double2 p15, p27, p36, p48;
      if (cz) {
        p15 = p27 = p36 = p48 = make_double2(0.0, 0.0);
      } else {
        const long j = pt.q0 + (k - nT);
        p15 = f.cstr[j];
        p27 = f.cstr[f.nk + j];
        p36 = f.cstr[2 * f.nk + j];
        p48 = f.cstr[3 * f.nk + j];
      }
      lds[k] = p15;
      lds[cap + k] = p27;
      lds[2 * cap + k] = p36;
      lds[3 * cap + k] = p48;
    }
  };
  // entries me + B, me + 2B, ... first and entry me (kept for level 2) last,
  // so no other evaluation overlaps the lifetime of its stresses
  for (int k = me + B; k < nE; k += B) level1(k);
  if (me < nE) level1(me);
  // level 1, stepu, entry me: inputs loaded before the barrier
  int q1 = 0;
  UIn u1;
  double2 v1;
  if (me < nU) {
    q1 = uslot(me);
    u1 = uload(me, q1);
    v1 = bf.vel[q1];
  }
  __syncthreads();
  // level 1, stepu: every U1 entry into the velocity table
  if (me < nU) {
    evp1d::StepuOut o;
    stepu(u1, v1, o);
    vt[me] = make_double2(o.uvel, o.vvel);
    if (TRAJ && me < nUo) out2<NT>(f.tr_mid + (q1 - d.u_begin), vt[me]);
  }
  for (int k = me + B; k < nU; k += B) {
    const int q = uslot(k);
    const UIn u = uload(k, q);
    evp1d::StepuOut o;
    stepu(u, bf.vel[q], o);
    vt[k] = make_double2(o.uvel, o.vvel);
    if (TRAJ && k < nUo) out2<NT>(f.tr_mid + (q - d.u_begin), vt[k]);
  }
  // level 2, stress of entry me: memory inputs loaded before the barrier
  // (velocities that are no U1 entry never change, kNoEntry: from memory)
  // (velocities that are no U1 entry never change, kNoEntry: from memory;
  // a ghost on the copy list whose source is no U1 entry: kFixBase + j, the
  // source's unchanging velocity or zero)
  constexpr unsigned kNo = evp1d::PairPlan::kNoEntry, kFix = evp1d::PairPlan::kFixBase;
  int g2 = 0;
  uint2 w2 = make_uint2(0u, 0u);
  double2 c2[4];
  if (me < nS) {
    const int4 n = f.tnbr[pt.n0 + me];
    g2 = me < nO ? tslot(me) : n.x;
    w2 = f.vloc[pt.v0 + me];
    c2[0] = bf.vel[g2];
    if (!FIX) {   // no copy list: every entry is a table entry or kNoEntry
      if ((w2.x >> 16) == kNo || (w2.y & 0xffffu) == kNo || (w2.y >> 16) == kNo) {
        c2[1] = bf.vel[n.y];
        c2[2] = bf.vel[n.z];
        c2[3] = bf.vel[n.w];
      }
    } else {
      const unsigned e[3] = {w2.x >> 16, w2.y & 0xffffu, w2.y >> 16};
      const int ns[3] = {n.y, n.z, n.w};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (e[c] == kNo) {
          c2[1 + c] = bf.vel[ns[c]];
        } else if (e[c] >= kFix) {
          const int s = f.fix[pt.f0 + (int)(e[c] - kFix)];
          c2[1 + c] = s < 0 ? make_double2(0.0, 0.0) : bf.vel[s];
        }
      }
    }
  }
  __syncthreads();
  // level 2, stress: owned and halo entries, one per thread
  if (me < nS) {
    const unsigned e0 = w2.x & 0xffffu, e1 = w2.x >> 16, e2 = w2.y & 0xffffu, e3 = w2.y >> 16;
    const double2 cc = e0 >= kFix ? c2[0] : vt[e0];
    const double2 ve = e1 >= kFix ? c2[1] : vt[e1];
    const double2 vn = e2 >= kFix ? c2[2] : vt[e2];
    const double2 vs = e3 >= kFix ? c2[3] : vt[e3];
    evp1d::StressOut o;
    evp1d::stress_point(d.s, cc.x, cc.y, ve.x, ve.y, vs.x, vs.y, vn.x, vn.y, m2[0].x, m2[0].y, st2, m2[1].x,
                        m2[1].y, m2[2].x, m2[2].y, sg, o);
    const int g = g2;
    if (me < nO) {
      out2<NT>(bf.sg_out + g, make_double2(o.sp[0], o.sp[1]));
      out2<NT>(bf.sg_out + (f.ls + g), make_double2(o.sp[2], o.sp[3]));
      out2<NT>(bf.sg_out + (2 * f.ls + g), make_double2(o.sm[0], o.sm[1]));
      out2<NT>(bf.sg_out + (3 * f.ls + g), make_double2(o.sm[2], o.sm[3]));
      out2<NT>(bf.sg_out + (4 * f.ls + g), make_double2(o.s12[0], o.s12[1]));
      out2<NT>(bf.sg_out + (5 * f.ls + g), make_double2(o.s12[2], o.s12[3]));
      if (last) {
        out2<NT>(d.s15 + g, make_double2(o.str[0], o.str[4]));
        out2<NT>(d.s27 + g, make_double2(o.str[1], o.str[6]));
        out2<NT>(d.s36 + g, make_double2(o.str[2], o.str[5]));
        out2<NT>(d.s48 + g, make_double2(o.str[3], o.str[7]));
      }
    }
    put(me, o);
  }
  // level 2, stepu of entry me: inputs loaded before the barrier
  int q2 = 0;
  UIn u2;
  if (me < nUo) {
    q2 = uslot(me);
    u2 = uload(me, q2);
  }
  __syncthreads();
  // level 2, stepu: the owned U points
  auto level2u = [&](int k, int q, const UIn &u) {
    evp1d::StepuOut o;
    stepu(u, vt[k], o);
    out2<NT>(bf.vel_out + q, make_double2(o.uvel, o.vvel));
    if (TRAJ && f.tr_end) out2<NT>(f.tr_end + (q - d.u_begin), make_double2(o.uvel, o.vvel));
    if (last) {
      out1<NT>(d.strintx + q, o.strintx);
      out1<NT>(d.strinty + q, o.strinty);
      out1<NT>(d.cb + q, o.Cb);
    }
  };
  if (me < nUo) level2u(me, q2, u2);
  for (int k = me + B; k < nUo; k += B) {
    const int q = uslot(k);
    level2u(k, q, uload(k, q));
  }
}

template <int B, bool X, bool TRAJ = false, bool FIX = false, bool NT = false>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(4))) void k_pair(DevView d, PairView f, int last) {
  pair_tile<B, TRAJ, FIX, NT>(d, f, PairBufs{d.vel, f.vel_out, f.sg_in, f.sg_out}, last,
                              tile_of(blockIdx.x, gridDim.x, X));
}

// The pair launches j0..L of a run (numbered as issue_fused numbers them) in
// one cooperative launch.  The workgroups form P groups (P = 8, one per XCD
// when the grid is a multiple of 8 and the tile order XCD-aware: workgroup b
// runs on XCD b mod 8; else P = 1), and group x runs the tiles that the
// one-launch form gives XCD x (tile_of).  Per launch, the group's workgroups
// take its tiles from a counter, so tiles of uneven cost balance out as they
// do when the hardware dispatches a launch; then a grid barrier, and the
// ghost copy with another barrier when a copy list is set.  The velocities
// alternate between d.vel and a.vy, the packed stresses between a.sg0 and
// a.sg1, exactly as the separate launches alternate them, so results are
// bit-identical.  a.cnt: 2 x 8 counters, zero at the launch, each on a
// 256-byte line of its own (the groups' atomics do not queue on one line);
// launch j uses row j mod 2 and clears the other row for launch j + 1.
constexpr int kCntPad = 64;   // ints from one tile counter to the next
struct PersistArgs {
  double2 *vy, *sg0, *sg1;
  int *cnt;
  int j0, L, ntiles;
};

template <int B, bool X, bool FIX>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(4))) void k_pair_persistent(DevView d, PairView f,
                                                                                             PersistArgs a) {
  cg::grid_group grid = cg::this_grid();
  // The arguments are read in place in the kernel-argument segment (d at
  // offset 0, f and a after it, each at its alignment), through a pointer
  // that an empty asm statement makes opaque once per tile: their fields are
  // then loaded where the tile body uses them, as in k_pair, instead of
  // being hoisted out of the loops and held in registers for the whole
  // launch (which spilled).
  typedef const __attribute__((address_space(4))) char *ArgP;
  constexpr size_t kOffF = (sizeof(DevView) + alignof(PairView) - 1) / alignof(PairView) * alignof(PairView);
  constexpr size_t kOffA =
      (kOffF + sizeof(PairView) + alignof(PersistArgs) - 1) / alignof(PersistArgs) * alignof(PersistArgs);
  ArgP ka = (ArgP)__builtin_amdgcn_kernarg_segment_ptr();
  const int j0 = a.j0, L = a.L, ntiles = a.ntiles;
  // this workgroup's group x and the group's tiles [t0, t0 + nt)
  const int P = (X && gridDim.x >= kXcds && gridDim.x % kXcds == 0) ? kXcds : 1;
  const int x = blockIdx.x % P;
  const int q = ntiles / P, r = ntiles % P;
  const int t0 = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q, nt = q + (x < r ? 1 : 0);
  __shared__ int grab;
  for (int j = j0; j <= L; ++j) {
    const bool to_x = ((L - j) % 2) == 0;   // the run ends in the handle's own buffer
    const bool odd = ((j - j0) % 2) == 1;   // which packed stress buffer is the input
    int *cnt = a.cnt + (j % 2) * kXcds * kCntPad;
    if (blockIdx.x == 0 && threadIdx.x < kXcds)   // with the atomics' scope, published by the barrier
      __hip_atomic_store(a.cnt + (((j + 1) % 2) * kXcds + threadIdx.x) * kCntPad, 0, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
      if (threadIdx.x == 0) grab = atomicAdd(cnt + x * kCntPad, 1);
      __syncthreads();
      const int k = grab;   // read by every thread before the tile's first barrier
      if (k >= nt) break;
      asm volatile("" : "+s"(ka));
      const DevView &dv = *(const DevView *)(const __attribute__((address_space(4))) DevView *)ka;
      const PairView &fv = *(const PairView *)(const __attribute__((address_space(4))) PairView *)(ka + kOffF);
      const PersistArgs &av =
          *(const PersistArgs *)(const __attribute__((address_space(4))) PersistArgs *)(ka + kOffA);
      const PairBufs bf{to_x ? av.vy : dv.vel, to_x ? dv.vel : av.vy, odd ? av.sg1 : av.sg0, odd ? av.sg0 : av.sg1};
      pair_tile<B, false, FIX, false>(dv, fv, bf, j == L ? 1 : 0, t0 + k);
      __syncthreads();   // the next tile rewrites the LDS tables and grab
    }
    grid.sync();
    if (d.nhalo > 0) {   // the copy after the launch's second subcycle
      double2 *out = to_x ? d.vel : a.vy;
      for (int k = blockIdx.x * B + threadIdx.x; k < d.nhalo; k += gridDim.x * B) {
        const int dst = d.halo_dst[k], src = d.halo_src[k];
        out[dst] = src >= 0 ? out[src] : make_double2(0.0, 0.0);
      }
      grid.sync();
    }
  }
}

// str1..8 of the constant entries, gathered once per run; *cnz (cleared by
// k_clear_flag first) becomes 1 if any of them has a bit set
__global__ void k_clear_flag(int *flag) { *flag = 0; }
// the shared points' stresses into the copy before the first subcycle, and
// back after the last
__global__ void k_scatter_shared(double *sig, long ldsig, const double *hsig, const int32_t *shared, long nsh) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsh) return;
  const long g = shared[j];
#pragma unroll
  for (int c = 0; c < 12; ++c) sig[c * ldsig + g] = hsig[j * kShStride + c];
}
__global__ void k_fill_shared(double *hsig, const double *sig, long ldsig, const int32_t *shared, long nsh) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsh) return;
  const long g = shared[j];
#pragma unroll
  for (int c = 0; c < 12; ++c) hsig[j * kShStride + c] = sig[c * ldsig + g];
}
__global__ void k_fill_const(DevView d, const int32_t *klist, double2 *cstr, long nk, int *cnz) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nk) return;
  const int g = klist[j];
  const double2 v[4] = {d.s15[g], d.s27[g], d.s36[g], d.s48[g]};
  unsigned long long bits = 0;
  for (int c = 0; c < 4; ++c) {
    cstr[c * nk + j] = v[c];
    bits |= (unsigned long long)__double_as_longlong(v[c].x) | (unsigned long long)__double_as_longlong(v[c].y);
  }
  if (bits) *cnz = 1;   // every writer stores the same value
}

constexpr int kFusedMaxEntries = 2560;   // 160 KiB of LDS per workgroup (gfx950)
constexpr int kFusedMinSub = 4;          // evp1d_run in AUTO mode: fused from this many subcycles
// pair form: block sizes (one level-2 stress per thread) and LDS limit
constexpr int kNumPairBlocks = 3;
constexpr int kPairBlocks[kNumPairBlocks] = {256, 512, 1024};
constexpr int kPairMaxLds = 160 * 1024;
// LDS of a pair launch: str table and velocity table
static size_t pair_lds(int cap, int capv) { return (size_t)(4 * cap + capv) * sizeof(double2); }

// launch configurations the autotuner chooses from
constexpr int kNumBlocks = 5;
constexpr int kBlocks[kNumBlocks] = {64, 128, 256, 512, 1024};
using StressFn = void (*)(DevView);
using StepuFn = void (*)(DevView, int);
static const StressFn kStressFns[kNumBlocks][2] = {
    {k_stress<64, false>, k_stress<64, true>},     {k_stress<128, false>, k_stress<128, true>},
    {k_stress<256, false>, k_stress<256, true>},   {k_stress<512, false>, k_stress<512, true>},
    {k_stress<1024, false>, k_stress<1024, true>}};
static const StepuFn kStepuFns[kNumBlocks][2] = {
    {k_stepu<64, false>, k_stepu<64, true>},       {k_stepu<128, false>, k_stepu<128, true>},
    {k_stepu<256, false>, k_stepu<256, true>},     {k_stepu<512, false>, k_stepu<512, true>},
    {k_stepu<1024, false>, k_stepu<1024, true>}};

struct LaunchCfg {
  int bi = 2;      // index into kBlocks
  int xcd = 0;
};

__global__ __launch_bounds__(kBlock) void k_halo(DevView d) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < d.nhalo) halo_at(d, k);
}

// Whole subcycle loop in one cooperative launch.  Every wave runs the same
// number of grid barriers (2 or 3 per subcycle) and leaves the loop together.
__global__ __launch_bounds__(kBlock) void k_persistent(DevView d, int nsub) {
  cg::grid_group grid = cg::this_grid();
  const int stride = gridDim.x * kBlock;
  const int t0 = blockIdx.x * kBlock + threadIdx.x;
  for (int k = 0; k < nsub; ++k) {
    for (int iw = t0; iw < d.t_end; iw += stride) stress_at(d, iw);
    grid.sync();
    const bool last = (k == nsub - 1);
    for (int iw = d.u_begin + t0; iw < d.u_end; iw += stride) stepu_at(d, iw, last);
    grid.sync();
    if (d.nhalo > 0) {
      for (int g = t0; g < d.nhalo; g += stride) halo_at(d, g);
      grid.sync();
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint kernels (reverse mode of one subcycle, see evp1d_adjoint.hpp)
// ---------------------------------------------------------------------------
struct AdjView {
  double2 *ub;            // [navel] adjoint of (uvel, vvel)
  double *sigb;           // [12][ldsig] adjoint of the stresses
  double2 *g;             // [navel] uarear * adjoint of (cc1, cc2); zero off the stepu range
  double2 *vc;            // [4][na] stress-adjoint velocity contributions: cc, ee, se, ne
  double *gstr, *gcw, *gfx, *gfy;   // [na] gradient accumulators
  const int *cnz;         // nonzero when str1..8 at some slot >= t_end (never written) has a bit set
  const double2 *pc;      // [kPairConst][lp]: the constant fields in pairs (k_pack_const)
  long lp;
  const int4 *nbr;        // per stress slot: ee, ne, se (k_pack_nbr)
};

// *cnz (zero on entry) becomes 1 if str1..8 at any slot in [t_end, navel)
// has a bit set; those slots' values never change during a run
__global__ void k_const_flag(DevView d, int *cnz) {
  const long j = d.t_end + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.navel) return;
  const double2 v[4] = {d.s15[j], d.s27[j], d.s36[j], d.s48[j]};
  unsigned long long bits = 0;
  for (int c = 0; c < 4; ++c)
    bits |= (unsigned long long)__double_as_longlong(v[c].x) | (unsigned long long)__double_as_longlong(v[c].y);
  if (bits) *cnz = 1;   // every writer stores the same value
}

// str1..8 of subcycle k from the state before it (stresses are not stored)
__global__ __launch_bounds__(256) void k_str_recompute(DevView d) {
  const int iw = blockIdx.x * 256 + threadIdx.x;
  if (iw >= d.t_end) return;
  const double2 cc = d.vel[iw];
  const double2 ve = d.vel[d.ee[iw]];
  const double2 vn = d.vel[d.ne[iw]];
  const double2 vs = d.vel[d.se[iw]];
  double sig[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) sig[c] = d.sig[c * d.ldsig + iw];
  evp1d::StressOut o;
  evp1d::stress_point(d.s, cc.x, cc.y, ve.x, ve.y, vs.x, vs.y, vn.x, vn.y, d.dxT[iw], d.dyT[iw],
                      d.strength[iw], d.hte[iw], d.htn[iw], d.htem1[iw], d.htnm1[iw], sig, o);
  d.s15[iw] = make_double2(o.str[0], o.str[4]);
  d.s27[iw] = make_double2(o.str[1], o.str[6]);
  d.s36[iw] = make_double2(o.str[2], o.str[5]);
  d.s48[iw] = make_double2(o.str[3], o.str[7]);
}

// Reverse of stepu for subcycle k.  The adjoint of u_k at q is the stepu
// part stored in ub (from subcycle k+1's stepu adjoint, or the seed) plus
// the contributions of subcycle k+1's stress adjoints that read q: its own
// (cc), and those of nw(q), sw(q), sse(q), which read q as their ee, ne, se
// neighbour (the inverse maps).  Only stepu slots need it: nothing else reads
// the velocity adjoint.  The update's result u_{k+1} is read from the stored
// trajectory (SPLIT), or recomputed from the str1..8 that k_str_recompute
// left in the working arrays: stepu_point is the forward kernels' own
// function, so both give the same bits.
//
// NT (with SPLIT): the cache hints (evp1d_cache.hpp) on the stored
// trajectory's velocities, the adjoint stresses, the velocity contributions
// and the gradient accumulators; the constant fields, ub and g stay cached.

// a velocity of subcycle k's start state: from the stored trajectory (split
// view, stepu slots) or the handle's arrays
template <bool SPLIT, bool NT>
__device__ __forceinline__ double2 adj_vel(const DevView &d, int j) {
  if (SPLIT && j >= d.u_begin && j < d.u_end) return in2<NT>(d.vel + (j - d.u_begin));
  return Ld<SPLIT>::vel(d, j);
}
template <bool NT>
__device__ __forceinline__ void adj_acc(double *p, double v) {
  out1<NT>(p, in1<NT>(p) + v);
}

template <bool SPLIT, bool NT = false>
__global__ __launch_bounds__(256) void k_stepu_adj(DevView d, AdjView a) {
  const int iw = d.u_begin + blockIdx.x * 256 + threadIdx.x;
  if (iw >= d.u_end) return;
  const int jw = d.nw[iw], js = d.sse[iw], jsw = d.sw[iw];
  const long na = d.na;
  double2 ub = a.ub[iw];
  if (iw < d.t_end) { const double2 t = in2<NT>(a.vc + iw); ub.x += t.x; ub.y += t.y; }
  if (jw < d.t_end) { const double2 t = in2<NT>(a.vc + (na + jw)); ub.x += t.x; ub.y += t.y; }
  if (jsw < d.t_end) { const double2 t = in2<NT>(a.vc + (3 * na + jsw)); ub.x += t.x; ub.y += t.y; }
  if (js < d.t_end) { const double2 t = in2<NT>(a.vc + (2 * na + js)); ub.x += t.x; ub.y += t.y; }
  const double2 v = adj_vel<SPLIT, NT>(d, iw);
  double2 vn;
  if (SPLIT) {
    vn = in2<NT>(d.vel_out + (iw - d.u_begin));
  } else {
    // str1..8 of slots stress never writes are +0.0 unless *cnz: no loads then
    const bool cz = *a.cnz == 0;
    auto str = [&](const double2 *p, int j) { return (cz && j >= d.t_end) ? make_double2(0.0, 0.0) : p[j]; };
    const double2 s15 = str(d.s15, iw), s27 = str(d.s27, jw), s36 = str(d.s36, js), s48 = str(d.s48, jsw);
    evp1d::StepuOut f;
    evp1d::stepu_point(d.s, v.x, v.y, d.cw[iw], d.aiu[iw], d.uocn[iw], d.vocn[iw], d.waterx[iw], d.watery[iw],
                       d.forcex[iw], d.forcey[iw], d.umassdti[iw], d.fm[iw], d.uarear[iw], d.uvel_init[iw],
                       d.vvel_init[iw], d.tbu[iw], s15.x, s27.x, s36.x, s48.x, s15.y, s36.y, s27.y, s48.y, f);
    vn = make_double2(f.uvel, f.vvel);
  }
  const long lp = a.lp;
  const double2 f0 = a.pc[3 * lp + iw], f1 = a.pc[4 * lp + iw], f2 = a.pc[5 * lp + iw];
  const double2 f4 = a.pc[7 * lp + iw], f5 = a.pc[8 * lp + iw];
  evp1d::StepuAdj o;
  evp1d::stepu_adjoint(d.s, v.x, v.y, f0.x, f0.y, f1.x, f1.y, f2.x, f2.y, f4.x, f4.y, f5.x, f5.y, vn.x, vn.y,
                       ub.x, ub.y, o);
  a.ub[iw] = make_double2(o.uold_b, o.vold_b);
  a.g[iw] = make_double2(o.gx, o.gy);
  adj_acc<NT>(a.gfx + iw, o.fx_b);
  adj_acc<NT>(a.gfy + iw, o.fy_b);
  adj_acc<NT>(a.gcw + iw, o.cw_b);
}

// Reverse of stress for subcycle k.  str_b of point q: its own terms and the
// terms stepu took from q as the nw, sse and sw neighbour of the points
// ee(q), se(q) and ne(q).
template <bool SPLIT, bool NT = false>
__global__ __launch_bounds__(256) void k_stress_adj(DevView d, AdjView a) {
  const int iw = blockIdx.x * 256 + threadIdx.x;
  if (iw >= d.t_end) return;
  const int4 nb = a.nbr[iw];
  const int ie = nb.x, in = nb.y, is = nb.z;
  const double2 cc = adj_vel<SPLIT, NT>(d, iw), ve = adj_vel<SPLIT, NT>(d, ie), vn = adj_vel<SPLIT, NT>(d, in),
                vs = adj_vel<SPLIT, NT>(d, is);
  const double2 g0 = a.g[iw], ge = a.g[ie], gs = a.g[is], gn = a.g[in];
  const double str_b[8] = {g0.x, ge.x, gs.x, gn.x, g0.y, gs.y, ge.y, gn.y};
  const long lp = a.lp;
  const double2 m0 = a.pc[iw], m1 = a.pc[lp + iw], m2 = a.pc[2 * lp + iw];
  double vb[8];
  const double gs_b = evp1d::stress_adjoint<NT>(d.s, cc.x, cc.y, ve.x, ve.y, vs.x, vs.y, vn.x, vn.y, m0.x, m0.y,
                                            d.strength[iw], m1.x, m1.y, m2.x, m2.y, a.sigb + iw, str_b,
                                            a.sigb + iw, d.ldsig, vb);
  const long na = d.na;
  out2<NT>(a.vc + iw, make_double2(vb[0], vb[1]));
  out2<NT>(a.vc + (na + iw), make_double2(vb[2], vb[3]));
  out2<NT>(a.vc + (2 * na + iw), make_double2(vb[4], vb[5]));
  out2<NT>(a.vc + (3 * na + iw), make_double2(vb[6], vb[7]));
  adj_acc<NT>(a.gstr + iw, gs_b);
}

// ---------------------------------------------------------------------------
// tangent-linear kernels: the forward point functions on Dual numbers
// ---------------------------------------------------------------------------
struct TanView {
  double *dsig;                         // [12][ldsig]
  double2 *dvel;                        // [navel]
  double2 *ds15, *ds27, *ds36, *ds48;   // [navel]
  const double *dstrength, *dcw, *dfx, *dfy;   // [na] direction
};

__device__ __forceinline__ evp1d::Dual D0(double v) { return evp1d::Dual{v, 0.0}; }
__device__ __forceinline__ evp1d::Dual DD(double v, double d) { return evp1d::Dual{v, d}; }

__global__ __launch_bounds__(256) void k_stress_tl(DevView d, TanView t) {
  using evp1d::Dual;
  const int iw = blockIdx.x * 256 + threadIdx.x;
  if (iw >= d.t_end) return;
  const int ie = d.ee[iw], in = d.ne[iw], is = d.se[iw];
  const double2 cc = d.vel[iw], ve = d.vel[ie], vn = d.vel[in], vs = d.vel[is];
  const double2 tc = t.dvel[iw], te = t.dvel[ie], tn = t.dvel[in], ts = t.dvel[is];
  Dual sig[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) sig[c] = DD(d.sig[c * d.ldsig + iw], t.dsig[c * d.ldsig + iw]);
  evp1d::StressOutT<Dual> o;
  evp1d::stress_point<Dual>(d.s, DD(cc.x, tc.x), DD(cc.y, tc.y), DD(ve.x, te.x), DD(ve.y, te.y),
                            DD(vs.x, ts.x), DD(vs.y, ts.y), DD(vn.x, tn.x), DD(vn.y, tn.y),
                            D0(d.dxT[iw]), D0(d.dyT[iw]), DD(d.strength[iw], t.dstrength[iw]),
                            D0(d.hte[iw]), D0(d.htn[iw]), D0(d.htem1[iw]), D0(d.htnm1[iw]), sig, o);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    d.sig[c * d.ldsig + iw] = o.sp[c].v;       t.dsig[c * d.ldsig + iw] = o.sp[c].d;
    d.sig[(4 + c) * d.ldsig + iw] = o.sm[c].v; t.dsig[(4 + c) * d.ldsig + iw] = o.sm[c].d;
    d.sig[(8 + c) * d.ldsig + iw] = o.s12[c].v; t.dsig[(8 + c) * d.ldsig + iw] = o.s12[c].d;
  }
  d.s15[iw] = make_double2(o.str[0].v, o.str[4].v); t.ds15[iw] = make_double2(o.str[0].d, o.str[4].d);
  d.s27[iw] = make_double2(o.str[1].v, o.str[6].v); t.ds27[iw] = make_double2(o.str[1].d, o.str[6].d);
  d.s36[iw] = make_double2(o.str[2].v, o.str[5].v); t.ds36[iw] = make_double2(o.str[2].d, o.str[5].d);
  d.s48[iw] = make_double2(o.str[3].v, o.str[7].v); t.ds48[iw] = make_double2(o.str[3].d, o.str[7].d);
}

__global__ __launch_bounds__(256) void k_stepu_tl(DevView d, TanView t, int write_diag) {
  using evp1d::Dual;
  const int iw = d.u_begin + blockIdx.x * 256 + threadIdx.x;
  if (iw >= d.u_end) return;
  const int jw = d.nw[iw], js = d.sse[iw], jsw = d.sw[iw];
  const double2 v = d.vel[iw], tv = t.dvel[iw];
  const double2 a = d.s15[iw], ta = t.ds15[iw];
  const double2 b = d.s27[jw], tb = t.ds27[jw];
  const double2 c = d.s36[js], tcs = t.ds36[js];
  const double2 e = d.s48[jsw], te = t.ds48[jsw];
  evp1d::StepuOutT<Dual> o;
  evp1d::stepu_point<Dual>(d.s, DD(v.x, tv.x), DD(v.y, tv.y), DD(d.cw[iw], t.dcw[iw]), D0(d.aiu[iw]),
                           D0(d.uocn[iw]), D0(d.vocn[iw]), D0(d.waterx[iw]), D0(d.watery[iw]),
                           DD(d.forcex[iw], t.dfx[iw]), DD(d.forcey[iw], t.dfy[iw]), D0(d.umassdti[iw]),
                           D0(d.fm[iw]), D0(d.uarear[iw]), D0(d.uvel_init[iw]), D0(d.vvel_init[iw]),
                           D0(d.tbu[iw]), DD(a.x, ta.x), DD(b.x, tb.x), DD(c.x, tcs.x), DD(e.x, te.x),
                           DD(a.y, ta.y), DD(c.y, tcs.y), DD(b.y, tb.y), DD(e.y, te.y), o);
  d.vel[iw] = make_double2(o.uvel.v, o.vvel.v);
  t.dvel[iw] = make_double2(o.uvel.d, o.vvel.d);
  if (write_diag) {
    d.strintx[iw] = o.strintx.v;
    d.strinty[iw] = o.strinty.v;
    d.cb[iw] = o.Cb.v;
  }
}

// loss and its directional derivative terms at the current state
__global__ void k_loss_tl(DevView d, TanView t, const double *ut, const double *vt, double *terms) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.na) return;
  double l = 0.0, dl = 0.0;
  if (q >= d.u_begin && q < d.u_end) {
    const double2 v = d.vel[q], tv = t.dvel[q];
    const double ex = v.x - ut[q], ey = v.y - vt[q];
    l = 0.5 * (ex * ex + ey * ey);
    dl = ex * tv.x + ey * tv.y;
  }
  terms[q] = l;
  terms[d.na + q] = dl;
}

// loss terms and adjoint seed at the final state (active U slots)
__global__ void k_seed(DevView d, AdjView a, const double *ut, const double *vt, double *terms) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.na) return;
  double t = 0.0;
  double2 s = make_double2(0.0, 0.0);
  if (q >= d.u_begin && q < d.u_end) {
    const double2 v = d.vel[q];
    s = make_double2(v.x - ut[q], v.y - vt[q]);
    t = 0.5 * (s.x * s.x + s.y * s.y);
  }
  if (a.ub) a.ub[q] = s;
  terms[q] = t;
}

// data movement between caller order and slot order.  Element t of a host
// array of member k sits at caller position q = a_off + t (t < m_na, active
// points) or g_off + t - m_na (ghost slots); caller position q < na lives in
// slot inv[q], ghost positions q >= na are their own slot.
struct Span {
  int a_off, m_na, g_off, n;
};
__device__ __forceinline__ int slot_of_t(const Span &sp, const int32_t *inv, int na, int t) {
  const int q = t < sp.m_na ? sp.a_off + t : sp.g_off + (t - sp.m_na);
  return q < na ? inv[q] : q;
}
__global__ void k_scatter_plain(double *dst, const double *src, const int32_t *inv, int na, Span sp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < sp.n) dst[slot_of_t(sp, inv, na, t)] = src[t];
}
__global__ void k_gather_plain(double *dst, const double *src, const int32_t *inv, int na, Span sp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < sp.n) dst[t] = src[slot_of_t(sp, inv, na, t)];
}
__global__ void k_scatter_comp(double2 *dst, const double *src, const int32_t *inv, int na, Span sp, int comp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < sp.n) {
    const int s = slot_of_t(sp, inv, na, t);
    if (comp == 0) dst[s].x = src[t]; else dst[s].y = src[t];
  }
}
__global__ void k_gather_comp(double *dst, const double2 *src, const int32_t *inv, int na, Span sp, int comp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < sp.n) {
    const int s = slot_of_t(sp, inv, na, t);
    dst[t] = (comp == 0) ? src[s].x : src[s].y;
  }
}
// relayout: new slot s takes old slot map[s] (s < na)
__global__ void k_permute_plain(double *dst, const double *src, const int32_t *map, int na) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < na) dst[s] = src[map[s]];
}
__global__ void k_permute_pair(double2 *dst, const double2 *src, const int32_t *map, int na) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < na) dst[s] = src[map[s]];
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
// 2D <-> 1D through the slot -> cell map (convert_2d_1d_dyn / convert_1d_2d_dyn)
__global__ void k_pack2d_plain(double *dst, const double *g, const int32_t *pos, int na) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < na) dst[s] = g[pos[s]];
}
__global__ void k_pack2d_comp(double2 *dst, const double *g, const int32_t *pos, int na, int navel, int comp,
                              int zero_ghosts) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= navel || (s >= na && !zero_ghosts)) return;
  const double v = s < na ? g[pos[s]] : 0.0;
  double *e = reinterpret_cast<double *>(dst + s);
  e[comp] = v;
}
__global__ void k_unpack2d_plain(double *g, const double *src, const int32_t *pos, int na) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < na) g[pos[s]] = src[s];
}
__global__ void k_unpack2d_comp(double *g, const double2 *src, const int32_t *pos, int na, int comp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < na) g[pos[s]] = reinterpret_cast<const double *>(src + s)[comp];
}

// neighbour list in slot space: slot s of point perm[s] gets the slot of its
// neighbour (ghost positions >= na are their own slot)
__global__ void k_idx_to_slots(int32_t *dst, const int32_t *src, const int32_t *perm, const int32_t *inv, int na) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= na) return;
  const int32_t q = src[perm[s]];
  dst[s] = q < na ? inv[q] : q;
}

struct evp1d_handle {
  int na = 0, navel = 0;
  long ldsig = 0;
  int device = 0;
  evp1d_params p{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void *> allocs;
  DevView v{};
  double *scratch = nullptr;       // navel doubles: staging for host copies
  double2 *ptmp = nullptr;         // na double2: relayout staging
  int32_t *idx_dev[6] = {};
  int32_t *inv_dev = nullptr;      // caller position -> slot
  int32_t *perm_dev = nullptr;     // slot -> caller position (slots < na)
  bool perm_dev_dirty = true;
  int32_t *idx_caller_dev[6] = {}; // neighbour lists in caller order (remapped to slots on the device)
  bool idx_caller_dirty[6] = {true, true, true, true, true, true};
  int32_t *map_dev = nullptr;      // relayout map
  int32_t *halo_dev = nullptr;     // 2*halo_cap: dst then src (slot space)
  int halo_cap = 0;
  // host mirrors in caller order (0-based)
  std::vector<int32_t> idx_host[6];
  bool idx_set[6] = {};
  std::vector<uint8_t> skip_host;  // bit0 skipT, bit1 skipU
  std::vector<int32_t> perm, inv;  // slot -> caller position, and inverse
  std::vector<int32_t> halo_dst_host, halo_src_host;   // caller order, -1 = fill
  bool layout_dirty = false;       // skip flags changed since the last relayout
  bool idx_dirty = false;          // slot-space lists need re-upload
  int mode = EVP1D_MODE_AUTO;
  LaunchCfg cfg_stress, cfg_stepu;
  // ensemble members: member k owns caller positions [a_off[k], a_off[k]+m_na[k])
  // and ghost positions [g_off[k], g_off[k]+m_navel[k]-m_na[k])
  std::vector<int32_t> m_na, m_navel, a_off, g_off;
  std::vector<uint8_t> idx_mset[6];   // per member: list set
  int coop_blocks = 0;
  int *pcnt_dev = nullptr;         // persistent pair form: tile counters (2 x 8, kCntPad apart)
  // fused form: tiling and its device tables (owned here, replaced on relayout)
  int tile_w = 64, tile_h = 16, tile_block = 512, tile_xcd = 1;
  bool tiling_auto = true;         // choose the tile from the domain size (until evp1d_set_tiling)
  evp1d::GridCoords coords;        // grid coordinates of the current lists
  bool coords_ok = false;
  evp1d::TileIds tile_cache;       // tile of every point for the current lists and tile size
  bool tile_cache_ok = false;
  bool fused_ok = false;
  std::string fused_why = "lists not set";
  int fused_ntiles = 0, fused_cap = 0;
  long nk_total = 0;
  evp1d::TileDesc *tiles_dev = nullptr;
  int32_t *hlist_dev = nullptr, *klist_dev = nullptr;
  int32_t *hid_dev = nullptr, *shared_dev = nullptr;
  long nsh = 0;                    // shared points (stress points some other tile evaluates as halo)
  double *hsig_dev = nullptr;      // their stresses, two buffers of nsh x kShStride
  uint2 *uloc_dev = nullptr;
  double2 *cstr_dev = nullptr;
  int *cnz_dev = nullptr;          // constant entries not all +0.0
  double2 *vel_alt = nullptr;      // second velocity buffer (navel)
  double *sig_alt = nullptr;       // second stress buffer (12 x ldsig), two-buffer fused runs
  int fused_inplace = 0;           // fused runs update stresses in place
  // pair form (two subcycles per fused launch): its tables, planned on the
  // current tiling when fused_depth is 2 (the default)
  int fused_depth = 2;
  int cache_hints = -1;            // pair form's cache hints: -1 chosen by size (pair_hints), 0 off, 1 on
  int tile_order = 0;              // 0 rows of tiles, 1 Hilbert curve over the tiles
  // subcycles per checkpointing unit of a gradient with fewer stored states
  // than subcycles (0: chosen per call)
  int adj_segment = 0;
  bool pair_ok = false;
  std::string pair_why = "pair form not selected";
  int pair_block = 0, pair_cap = 0, pair_capv = 0;
  long pk_total = 0;
  evp1d::PairTile *ptiles_dev = nullptr;
  int32_t *pxlist_dev = nullptr, *pelist_dev = nullptr, *pklist_dev = nullptr;
  uint2 *puloc_dev = nullptr, *pvloc_dev = nullptr;
  int4 *ptnbr_dev = nullptr;
  int32_t *pfix_dev = nullptr;
  double2 *pcstr_dev = nullptr;
  double2 *pconst = nullptr;       // constant fields in pairs, [kPairConst][na], packed per run
  double *sig_pair = nullptr;      // stresses in pairs during a pair run (with sig_alt: two buffers)
  int4 *nbr4 = nullptr;            // ee, ne, se per stress slot, packed per pair run
  void *traj = nullptr;            // gradient trajectory storage, kept between calls
  size_t traj_bytes = 0;
  void *adj_ws = nullptr;          // the gradient's per-call buffers, kept between calls
  size_t adj_ws_bytes = 0;
  void *ckpt = nullptr;            // a checkpointed gradient's stored states, kept between calls
  size_t ckpt_bytes = 0;
  bool pconst_fresh = false;       // pconst holds the current fields (set for a gradient call's duration)
  bool fills_fresh = false;        // str1..8 of the constant entries and *cnz filled by this gradient call
  // 2D grid of the caller's arrays (device-side pack/unpack)
  int grid_nx = 0, grid_ny = 0;
  std::vector<int32_t> grid_ij;    // caller position < na -> 0-based linear 2D index
  int32_t *pos2d_dev = nullptr;    // slot < na -> linear 2D index in the current slot order
  bool pos2d_dirty = true;
  double *g2d[EVP1D_NFIELDS] = {}; // device copy of each field's 2D array
  int graph_mode = -1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_nsub = -1;
  bool graph_dirty = true;
};

static Scalars scalars_from(const evp1d_params &p) {
  Scalars s;
  s.capping = p.capping; s.e_factor = p.e_factor; s.epp2i = p.epp2i;
  s.deltaminEVP = p.deltaminEVP; s.arlx1i = p.arlx1i; s.denom1 = p.denom1;
  s.brlx = p.brlx; s.revp = p.revp; s.Ktens = p.Ktens; s.u0 = p.u0;
  s.cosw = p.cosw; s.sinw = p.sinw; s.rhow = p.rhow;
  return s;
}

static int dev_alloc(evp1d_handle *h, void **ptr, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess)
    return fail(EVP1D_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  h->allocs.push_back(*ptr);
  HIP_TRY(hipMemsetAsync(*ptr, 0, bytes, h->stream));
  return EVP1D_OK;
}

static void drop_graph(evp1d_handle *h) {
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->graph_exec = nullptr;
  h->graph = nullptr;
  h->graph_nsub = -1;
}

static void free_pairs(evp1d_handle *h) {
  void *bufs[9] = {h->ptiles_dev, h->pxlist_dev, h->pelist_dev, h->pklist_dev, h->puloc_dev,
                   h->pvloc_dev,  h->pcstr_dev,  h->ptnbr_dev,  h->pfix_dev};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  h->ptiles_dev = nullptr;
  h->pxlist_dev = h->pelist_dev = h->pklist_dev = nullptr;
  h->puloc_dev = h->pvloc_dev = nullptr;
  h->ptnbr_dev = nullptr;
  h->pfix_dev = nullptr;
  h->pcstr_dev = nullptr;
  h->pair_ok = false;
  h->pair_block = h->pair_cap = h->pair_capv = 0;
  h->pk_total = 0;
}

static void free_tiling(evp1d_handle *h) {
  void *bufs[8] = {h->tiles_dev, h->hlist_dev, h->klist_dev, h->uloc_dev,
                   h->cstr_dev,  h->hid_dev,   h->shared_dev, h->hsig_dev};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  free_pairs(h);
  h->tiles_dev = nullptr;
  h->hlist_dev = h->klist_dev = nullptr;
  h->hid_dev = h->shared_dev = nullptr;
  h->hsig_dev = nullptr;
  h->nsh = 0;
  h->uloc_dev = nullptr;
  h->cstr_dev = nullptr;
  h->fused_ok = false;
  h->fused_ntiles = h->fused_cap = 0;
  h->nk_total = 0;
}

static void release(evp1d_handle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph(h);
  free_tiling(h);
  if (h->vel_alt) (void)hipFree(h->vel_alt);
  if (h->sig_alt) (void)hipFree(h->sig_alt);
  if (h->pconst) (void)hipFree(h->pconst);
  if (h->sig_pair) (void)hipFree(h->sig_pair);
  if (h->nbr4) (void)hipFree(h->nbr4);
  if (h->pcnt_dev) (void)hipFree(h->pcnt_dev);
  if (h->cnz_dev) (void)hipFree(h->cnz_dev);
  if (h->traj) (void)hipFree(h->traj);
  if (h->adj_ws) (void)hipFree(h->adj_ws);
  if (h->ckpt) (void)hipFree(h->ckpt);
  if (h->pos2d_dev) (void)hipFree(h->pos2d_dev);
  for (double *g : h->g2d)
    if (g) (void)hipFree(g);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

static inline int blocks_for(int n) { return (n + kBlock - 1) / kBlock; }

// field table ---------------------------------------------------------------
enum FieldKind { FK_PLAIN, FK_PAIR };
struct FieldRef {
  FieldKind kind;
  double *plain;   // FK_PLAIN (na doubles, slot order)
  double2 *pair;   // FK_PAIR  (navel double2, slot order)
  int comp;        // component of pair
  int len;         // caller-visible length
};

static bool field_ref(evp1d_handle *h, int f, FieldRef &r) {
  DevView &v = h->v;
  r = FieldRef{FK_PLAIN, nullptr, nullptr, 0, h->na};
  auto plain = [&](const double *p) { r.kind = FK_PLAIN; r.plain = const_cast<double *>(p); };
  auto pair = [&](double2 *p, int c) { r.kind = FK_PAIR; r.pair = p; r.comp = c; r.len = h->navel; };
  if (f >= EVP1D_STRESSP_1 && f <= EVP1D_STRESS12_4) {
    plain(v.sig + (long)(f - EVP1D_STRESSP_1) * v.ldsig);
    return true;
  }
  switch (f) {
    case EVP1D_STRENGTH: plain(v.strength); return true;
    case EVP1D_UVEL: pair(v.vel, 0); return true;
    case EVP1D_VVEL: pair(v.vel, 1); return true;
    case EVP1D_DXT: plain(v.dxT); return true;
    case EVP1D_DYT: plain(v.dyT); return true;
    case EVP1D_UAREAR: plain(v.uarear); return true;
    case EVP1D_CDN_OCN: plain(v.cw); return true;
    case EVP1D_AIU: plain(v.aiu); return true;
    case EVP1D_UOCN: plain(v.uocn); return true;
    case EVP1D_VOCN: plain(v.vocn); return true;
    case EVP1D_WATERX: plain(v.waterx); return true;
    case EVP1D_WATERY: plain(v.watery); return true;
    case EVP1D_FORCEX: plain(v.forcex); return true;
    case EVP1D_FORCEY: plain(v.forcey); return true;
    case EVP1D_UMASSDTI: plain(v.umassdti); return true;
    case EVP1D_FM: plain(v.fm); return true;
    case EVP1D_STRINTX: plain(v.strintx); return true;
    case EVP1D_STRINTY: plain(v.strinty); return true;
    case EVP1D_TBU: plain(v.tbu); return true;
    case EVP1D_HTE: plain(v.hte); return true;
    case EVP1D_HTN: plain(v.htn); return true;
    case EVP1D_HTEM1: plain(v.htem1); return true;
    case EVP1D_HTNM1: plain(v.htnm1); return true;
    case EVP1D_UVEL_INIT: plain(v.uvel_init); return true;
    case EVP1D_VVEL_INIT: plain(v.vvel_init); return true;
    case EVP1D_CB: plain(v.cb); return true;
    case EVP1D_STR1: pair(v.s15, 0); return true;
    case EVP1D_STR5: pair(v.s15, 1); return true;
    case EVP1D_STR2: pair(v.s27, 0); return true;
    case EVP1D_STR7: pair(v.s27, 1); return true;
    case EVP1D_STR3: pair(v.s36, 0); return true;
    case EVP1D_STR6: pair(v.s36, 1); return true;
    case EVP1D_STR4: pair(v.s48, 0); return true;
    case EVP1D_STR8: pair(v.s48, 1); return true;
    default: return false;
  }
}

static bool ready(const evp1d_handle *h) {
  for (int k = 0; k < 6; ++k)
    if (!h->idx_set[k]) return false;
  return true;
}

// caller-position span of member k (k < 0: the whole handle)
struct MemberSpan {
  int a_off, m_na, g_off, m_navel;
};
static bool member_span(const evp1d_handle *h, int k, MemberSpan &m) {
  if (k < 0) {
    m = MemberSpan{0, h->na, h->na, h->navel};
    return true;
  }
  if (k >= (int)h->m_na.size()) return false;
  m = MemberSpan{h->a_off[(size_t)k], h->m_na[(size_t)k], h->g_off[(size_t)k], h->m_navel[(size_t)k]};
  return true;
}

// every na-length device array (slot order), for relayout
static void plain_arrays(evp1d_handle *h, std::vector<double *> &out) {
  DevView &v = h->v;
  out = {const_cast<double *>(v.dxT), const_cast<double *>(v.dyT), const_cast<double *>(v.strength),
         const_cast<double *>(v.hte), const_cast<double *>(v.htn), const_cast<double *>(v.htem1),
         const_cast<double *>(v.htnm1), const_cast<double *>(v.cw), const_cast<double *>(v.aiu),
         const_cast<double *>(v.uocn), const_cast<double *>(v.vocn), const_cast<double *>(v.waterx),
         const_cast<double *>(v.watery), const_cast<double *>(v.forcex), const_cast<double *>(v.forcey),
         const_cast<double *>(v.umassdti), const_cast<double *>(v.fm), const_cast<double *>(v.uarear),
         const_cast<double *>(v.uvel_init), const_cast<double *>(v.vvel_init), const_cast<double *>(v.tbu),
         v.strintx, v.strinty, v.cb};
  for (int c = 0; c < 12; ++c) out.push_back(v.sig + (long)c * v.ldsig);
}

// Upload the fused form's tables for a plan made on the current slot order.
template <class T>
static int upload_table(evp1d_handle *h, T **dst, const T *src, size_t n) {
  HIP_TRY(hipMalloc((void **)dst, std::max<size_t>(n, 1) * sizeof(T)));
  if (n) HIP_TRY(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
  return EVP1D_OK;
}

static int install_tiling(evp1d_handle *h, const evp1d::TilePlan &plan) {
  HIP_TRY(hipStreamSynchronize(h->stream));   // no launch may still read the old tables
  drop_graph(h);
  free_tiling(h);
  h->fused_why = plan.why;
  if (!plan.ok) return EVP1D_OK;
  RC_TRY(upload_table(h, &h->tiles_dev, plan.tiles.data(), plan.tiles.size()));
  RC_TRY(upload_table(h, &h->hlist_dev, plan.list.data(), plan.list.size()));
  RC_TRY(upload_table(h, &h->klist_dev, plan.klist.data(), plan.klist.size()));
  RC_TRY(upload_table(h, &h->hid_dev, plan.hid.data(), plan.hid.size()));
  RC_TRY(upload_table(h, &h->shared_dev, plan.shared.data(), plan.shared.size()));
  h->nsh = (long)plan.shared.size();
  HIP_TRY(hipMalloc((void **)&h->hsig_dev, std::max<size_t>((size_t)2 * kShStride * h->nsh, 1) * sizeof(double)));
  static_assert(sizeof(uint2) == 4 * sizeof(uint16_t), "uloc packing");
  RC_TRY(upload_table(h, &h->uloc_dev, reinterpret_cast<const uint2 *>(plan.uloc.data()), plan.uloc.size() / 4));
  h->nk_total = (long)plan.klist.size();
  HIP_TRY(hipMalloc((void **)&h->cstr_dev, std::max<size_t>((size_t)4 * h->nk_total, 1) * sizeof(double2)));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->fused_ntiles = (int)plan.tiles.size();
  h->fused_cap = std::max(plan.max_entries, 1);
  h->fused_ok = true;
  return EVP1D_OK;
}

// The pair form's tables on the tiling just installed (when selected).
static int install_pairs(evp1d_handle *h, const evp1d::TilePlan &plan, const int32_t *const idx[6]) {
  free_pairs(h);
  if (h->fused_depth != 2) {
    h->pair_why = "pair form not selected";
    return EVP1D_OK;
  }
  // the ghost copy list in slot space (ghost slots are their own slots)
  std::vector<int32_t> hsrc(h->halo_src_host.size());
  for (size_t k = 0; k < hsrc.size(); ++k) {
    const int32_t s = h->halo_src_host[k];
    hsrc[k] = s < 0 ? -1 : h->inv[(size_t)s];
  }
  evp1d::HaloList hl;
  hl.n = (int)hsrc.size();
  hl.dst = h->halo_dst_host.data();
  hl.src = hsrc.data();
  const evp1d::PairPlan q = evp1d::plan_pairs(plan, h->na, h->navel, idx, h->skip_host.data(), kPairMaxLds / 64,
                                              kPairMaxLds / 16, kPairBlocks[kNumPairBlocks - 1], hl);
  if (!q.ok) {
    h->pair_why = q.why;
    return EVP1D_OK;
  }
  const std::string bad = evp1d::check_pairs(q, plan, h->na, h->navel, idx, h->skip_host.data(), hl);
  if (!bad.empty()) return fail(EVP1D_EINDEX, "pair plan: %s", bad.c_str());
  int b = 0;
  while (kPairBlocks[b] < q.max_t2) ++b;   // max_t2 <= the largest block (plan limit)
  const size_t lds = pair_lds(std::max(q.cap_s, 1), std::max(q.cap_v, 1));
  if (lds > (size_t)kPairMaxLds) {
    h->pair_why = "tables need " + std::to_string(lds) + " bytes of LDS";
    return EVP1D_OK;
  }
  RC_TRY(upload_table(h, &h->ptiles_dev, q.tiles.data(), q.tiles.size()));
  RC_TRY(upload_table(h, &h->pxlist_dev, q.xlist.data(), q.xlist.size()));
  RC_TRY(upload_table(h, &h->pelist_dev, q.elist.data(), q.elist.size()));
  RC_TRY(upload_table(h, &h->pklist_dev, q.klist.data(), q.klist.size()));
  RC_TRY(upload_table(h, &h->puloc_dev, reinterpret_cast<const uint2 *>(q.uloc.data()), q.uloc.size() / 4));
  RC_TRY(upload_table(h, &h->pvloc_dev, reinterpret_cast<const uint2 *>(q.vloc.data()), q.vloc.size() / 4));
  RC_TRY(upload_table(h, &h->ptnbr_dev, reinterpret_cast<const int4 *>(q.tnbr.data()), q.tnbr.size() / 4));
  RC_TRY(upload_table(h, &h->pfix_dev, q.fix.data(), q.fix.size()));
  h->pk_total = (long)q.klist.size();
  HIP_TRY(hipMalloc((void **)&h->pcstr_dev, std::max<size_t>((size_t)4 * h->pk_total, 1) * sizeof(double2)));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->pair_block = kPairBlocks[b];
  h->pair_cap = std::max(q.cap_s, 1);
  h->pair_capv = std::max(q.cap_v, 1);
  h->pair_ok = true;
  h->pair_why.clear();
  return EVP1D_OK;
}

// Rebuild the slot order (class, then fused tile, then caller position) from
// the skip flags and the neighbour lists, and move every resident array into
// it.  Without a complete set of lists the order is class, caller position.
static int relayout_impl(evp1d_handle *h);
static int relayout(evp1d_handle *h) {
  return evp1d::no_throw("slot layout", [h] { return relayout_impl(h); });
}
static int relayout_impl(evp1d_handle *h) {
  const int na = h->na;
  const int32_t *idx[6];
  bool all = na > 0;
  for (int k = 0; k < 6; ++k) {
    all = all && h->idx_set[k] && h->idx_host[k].size() == (size_t)na;
    idx[k] = all ? h->idx_host[k].data() : nullptr;
  }
  // the tile of each point depends on the lists and the tile size only: it
  // is kept while only skip flags change (a new model step's ice masks)
  if (!h->coords_ok) {
    h->coords = evp1d::grid_coords(na, h->navel, all ? idx : nullptr);
    h->coords_ok = true;
    h->tile_cache_ok = false;
  }
  if (h->tiling_auto && !h->tile_cache_ok) {
    // one subcycle per launch: enough tiles to fill the GPU, the largest
    // tile of the ladder that still gives min_tiles tiles (256 CUs); the
    // pair form: 31x19 cells, the tuner's choice from 84k to 5.4M points
    // (at most 640 level-2 stress points per tile: one 512- or 1024-thread
    // workgroup; one subcycle per launch on the same tiles where the pair
    // tables do not fit)
    struct Rung { int w, h, block, min_tiles; };
    static const Rung ladder[3] = {{64, 16, 512, 1024}, {32, 8, 256, 512}, {16, 8, 128, 0}};
    static const Rung pair_rung = {31, 19, 512, 0};
    const Rung *rungs = h->fused_depth == 2 ? &pair_rung : ladder;
    const int nrungs = h->fused_depth == 2 ? 1 : 3;
    for (int k = 0; k < nrungs; ++k) {
      const Rung &r = rungs[k];
      h->tile_cache = evp1d::tile_ids(h->coords, r.w, r.h, h->tile_order);
      h->tile_w = r.w;
      h->tile_h = r.h;
      h->tile_block = r.block;
      if (h->tile_cache.ntiles >= r.min_tiles) break;
    }
    h->tile_cache_ok = true;
  }
  if (!h->tile_cache_ok || h->tile_cache.tile_w != h->tile_w || h->tile_cache.tile_h != h->tile_h ||
      h->tile_cache.order != h->tile_order) {
    h->tile_cache = evp1d::tile_ids(h->coords, h->tile_w, h->tile_h, h->tile_order);
    h->tile_cache_ok = true;
  }
  const evp1d::TilePlan plan = evp1d::plan_tiles(na, h->navel, all ? idx : nullptr, h->skip_host.data(),
                                                 h->tile_cache, kFusedMaxEntries);
  const std::string bad = evp1d::check_plan(plan, na, h->navel, idx, h->skip_host.data());
  if (!bad.empty()) return fail(EVP1D_EINDEX, "slot layout: %s", bad.c_str());
  const std::vector<int32_t> &nperm = plan.perm;
  if (na > 0 && nperm != h->perm) {
    std::vector<int32_t> map((size_t)na), ninv((size_t)na);
    for (int s = 0; s < na; ++s) {
      map[(size_t)s] = h->inv[(size_t)nperm[(size_t)s]];
      ninv[(size_t)nperm[(size_t)s]] = s;
    }
    HIP_TRY(hipMemcpyAsync(h->map_dev, map.data(), (size_t)na * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    std::vector<double *> arrs;
    plain_arrays(h, arrs);
    double *tmp = reinterpret_cast<double *>(h->ptmp);
    for (double *a : arrs) {
      hipLaunchKernelGGL(k_permute_plain, dim3(blocks_for(na)), dim3(kBlock), 0, h->stream, tmp, a, h->map_dev, na);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(a, tmp, (size_t)na * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    }
    double2 *pairs[5] = {h->v.vel, h->v.s15, h->v.s27, h->v.s36, h->v.s48};
    for (double2 *a : pairs) {
      hipLaunchKernelGGL(k_permute_pair, dim3(blocks_for(na)), dim3(kBlock), 0, h->stream, h->ptmp, a, h->map_dev, na);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(a, h->ptmp, (size_t)na * sizeof(double2), hipMemcpyDeviceToDevice, h->stream));
    }
    HIP_TRY(hipMemcpyAsync(h->inv_dev, ninv.data(), (size_t)na * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->perm = nperm;
    h->inv.swap(ninv);
    h->pos2d_dirty = true;
    h->perm_dev_dirty = true;
  }
  RC_TRY(install_tiling(h, plan));
  RC_TRY(install_pairs(h, plan, all ? idx : nullptr));
  h->v.t_end = plan.cnt[0] + plan.cnt[1];
  h->v.u_begin = plan.cnt[0];
  h->v.u_end = plan.cnt[0] + plan.cnt[1] + plan.cnt[2];
  h->layout_dirty = false;
  h->idx_dirty = true;
  h->graph_dirty = true;
  return EVP1D_OK;
}

static inline int32_t slot_of(const evp1d_handle *h, int32_t q) { return q < h->na ? h->inv[(size_t)q] : q; }

// bring the device copy of layout, lists and halo up to date
static int ensure(evp1d_handle *h) {
  if (h->layout_dirty) RC_TRY(relayout(h));
  if (!h->idx_dirty) return EVP1D_OK;
  const int na = h->na;
  if (na > 0 && h->perm_dev_dirty) {
    HIP_TRY(hipMemcpyAsync(h->perm_dev, h->perm.data(), (size_t)na * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    h->perm_dev_dirty = false;
  }
  for (int k = 0; k < 6; ++k) {
    if (!h->idx_set[k] || na == 0) continue;
    if (h->idx_caller_dirty[k]) {   // uploaded once per evp1d_set_index
      HIP_TRY(hipMemcpyAsync(h->idx_caller_dev[k], h->idx_host[k].data(), (size_t)na * sizeof(int32_t),
                             hipMemcpyHostToDevice, h->stream));
      h->idx_caller_dirty[k] = false;
    }
    hipLaunchKernelGGL(k_idx_to_slots, dim3(blocks_for(na)), dim3(kBlock), 0, h->stream, h->idx_dev[k],
                       (const int32_t *)h->idx_caller_dev[k], (const int32_t *)h->perm_dev,
                       (const int32_t *)h->inv_dev, na);
    HIP_TRY(hipGetLastError());
  }
  if (na > 0) HIP_TRY(hipStreamSynchronize(h->stream));
  const int n = (int)h->halo_dst_host.size();
  if (n > 0) {
    std::vector<int32_t> buf((size_t)2 * n);
    for (int k = 0; k < n; ++k) {
      buf[(size_t)k] = h->halo_dst_host[(size_t)k];
      const int32_t src = h->halo_src_host[(size_t)k];
      buf[(size_t)n + k] = src < 0 ? -1 : slot_of(h, src);
    }
    HIP_TRY(hipMemcpyAsync(h->halo_dev, buf.data(), (size_t)2 * n * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  h->idx_dirty = false;
  h->graph_dirty = true;
  return EVP1D_OK;
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
static int launch_stress(evp1d_handle *h) {
  if (h->v.t_end == 0) return EVP1D_OK;
  const LaunchCfg c = h->cfg_stress;
  const int B = kBlocks[c.bi];
  hipLaunchKernelGGL(kStressFns[c.bi][c.xcd], dim3((h->v.t_end + B - 1) / B), dim3(B), 0, h->stream, h->v);
  HIP_TRY(hipGetLastError());
  return EVP1D_OK;
}

static int launch_stepu(evp1d_handle *h, bool write_diag) {
  const int n = h->v.u_end - h->v.u_begin;
  if (n == 0) return EVP1D_OK;
  const LaunchCfg c = h->cfg_stepu;
  const int B = kBlocks[c.bi];
  hipLaunchKernelGGL(kStepuFns[c.bi][c.xcd], dim3((n + B - 1) / B), dim3(B), 0, h->stream, h->v,
                     write_diag ? 1 : 0);
  HIP_TRY(hipGetLastError());
  return EVP1D_OK;
}

static int launch_halo(evp1d_handle *h) {
  if (h->v.nhalo == 0) return EVP1D_OK;
  hipLaunchKernelGGL(k_halo, dim3(blocks_for(h->v.nhalo)), dim3(kBlock), 0, h->stream, h->v);
  HIP_TRY(hipGetLastError());
  return EVP1D_OK;
}

static int issue_subcycles(evp1d_handle *h, int nsub) {
  for (int k = 0; k < nsub; ++k) {
    RC_TRY(launch_stress(h));
    RC_TRY(launch_stepu(h, k == nsub - 1));
    RC_TRY(launch_halo(h));
  }
  return EVP1D_OK;
}

typedef void (*FusedFn)(DevView, FusedView, int);
static const int kFusedBlocks[4] = {64, 128, 256, 512};
static const FusedFn kFusedFns[4][2] = {{k_fused<64, false>, k_fused<64, true>},
                                        {k_fused<128, false>, k_fused<128, true>},
                                        {k_fused<256, false>, k_fused<256, true>},
                                        {k_fused<512, false>, k_fused<512, true>}};
static const FusedFn kFusedSplitFns[4] = {k_fused<64, true, true>, k_fused<128, true, true>,
                                         k_fused<256, true, true>, k_fused<512, true, true>};
typedef void (*PairFn)(DevView, PairView, int);
// [workgroup size][XCD-aware order][cache hints]
#define PAIR_ROW(B, FIX)                                                                      \
  {{k_pair<B, false, false, FIX, false>, k_pair<B, false, false, FIX, true>},                 \
   {k_pair<B, true, false, FIX, false>, k_pair<B, true, false, FIX, true>}}
static const PairFn kPairFns[kNumPairBlocks][2][2] = {PAIR_ROW(256, false), PAIR_ROW(512, false),
                                                      PAIR_ROW(1024, false)};
// with a ghost copy list (level-2 reads of listed ghosts through PairPlan::fix)
static const PairFn kPairFixFns[kNumPairBlocks][2][2] = {PAIR_ROW(256, true), PAIR_ROW(512, true),
                                                         PAIR_ROW(1024, true)};
#undef PAIR_ROW
// the gradient sweep's trajectory-storing form (XCD-aware tile order only),
// [workgroup size][cache hints]
static const PairFn kPairTrajFns[kNumPairBlocks][2] = {
    {k_pair<256, true, true, false, false>, k_pair<256, true, true, false, true>},
    {k_pair<512, true, true, false, false>, k_pair<512, true, true, false, true>},
    {k_pair<1024, true, true, false, false>, k_pair<1024, true, true, false, true>}};
// one cooperative launch for the pair launches of a run (evp1d_run in
// EVP1D_MODE_PERSISTENT with the pair form selected)
typedef void (*PairPersistFn)(DevView, PairView, PersistArgs);
static const PairPersistFn kPairPersistFns[kNumPairBlocks][2][2] = {
    {{k_pair_persistent<256, false, false>, k_pair_persistent<256, false, true>},
     {k_pair_persistent<256, true, false>, k_pair_persistent<256, true, true>}},
    {{k_pair_persistent<512, false, false>, k_pair_persistent<512, false, true>},
     {k_pair_persistent<512, true, false>, k_pair_persistent<512, true, true>}},
    {{k_pair_persistent<1024, false, false>, k_pair_persistent<1024, false, true>},
     {k_pair_persistent<1024, true, false>, k_pair_persistent<1024, true, true>}}};
static int pair_block_index(int b) {
  for (int i = 0; i < kNumPairBlocks; ++i)
    if (kPairBlocks[i] == b) return i;
  return -1;
}
// the pair form runs when selected and planned
static bool pair_form(const evp1d_handle *h) { return h->fused_depth == 2 && h->pair_ok; }

// Whether the pair form's launches use the cache hints (NT).  By default:
// when all the bytes a launch touches exceed the last-level cache (256 MiB
// on MI355X) but the part every launch reads again (the U points' constant
// fields and tables, about 152 B per U point, and the velocities) is at most
// 1.3 times it (hints_pay).  Measured on the pair form at 31x19: 8 % less time per
// subcycle at 1.36M and 1.68M points (reread part 221 and 273 MB); 9-13 %
// more at 0.34M and 0.63M (everything fits: 113 and 211 MB), 4-7 % more at
// 2.5M and 5.4M (reread part 409 MB and more).
static bool pair_hints(const evp1d_handle *h) {
  if (h->cache_hints >= 0) return h->cache_hints != 0;
  const double nT = h->v.t_end, nU = h->v.u_end - h->v.u_begin;
  const double reread = 152.0 * nU + 16.0 * h->navel;
  return hints_pay(reread, reread + 248.0 * nT);   // + stresses in and out, the T points' constant fields
}

static int fused_block_index(int b) {
  for (int i = 0; i < 4; ++i)
    if (kFusedBlocks[i] == b) return i;
  return -1;
}

static int ensure_fused_buffers(evp1d_handle *h) {
  if (!h->cnz_dev) {
    HIP_TRY(hipMalloc((void **)&h->cnz_dev, 256));
    // tables above 64 KiB of LDS
    for (auto &row : kFusedFns)
      for (FusedFn fn : row)
        HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    kFusedMaxEntries * 4 * (int)sizeof(double2)));
    for (FusedFn fn : kFusedSplitFns)
      HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  kFusedMaxEntries * 4 * (int)sizeof(double2)));
    for (const auto *tab : {&kPairFns, &kPairFixFns})
      for (const auto &row : *tab)
        for (const auto &pair : row)
          for (PairFn fn : pair)
            HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, kPairMaxLds));
    for (const auto &row : kPairTrajFns)
      for (PairFn fn : row)
        HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, kPairMaxLds));
  }
  if (!h->vel_alt) HIP_TRY(hipMalloc((void **)&h->vel_alt, (size_t)h->navel * sizeof(double2)));
  if ((!h->fused_inplace || pair_form(h)) && !h->sig_alt)
    HIP_TRY(hipMalloc((void **)&h->sig_alt, (size_t)12 * h->ldsig * sizeof(double)));
  if (pair_form(h) && !h->pconst)
    HIP_TRY(hipMalloc((void **)&h->pconst, (size_t)kPairConst * std::max(h->na, 1) * sizeof(double2)));
  if (pair_form(h) && !h->sig_pair)
    HIP_TRY(hipMalloc((void **)&h->sig_pair, (size_t)12 * h->ldsig * sizeof(double)));
  return EVP1D_OK;
}


// Before the first fused subcycle of a run: the constant entries' str1..8
// and, for an in-place run, the shared points' stresses (into copy 0).
// Fills the run's view; fused_step points it at subcycle k.
static int fused_prologue(evp1d_handle *h, const DevView &d, FusedView &f, bool ip, bool fills = true) {
  if (fills) {
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, h->stream, h->cnz_dev);
    HIP_TRY(hipGetLastError());
  }
  if (fills && h->nk_total > 0) {
    hipLaunchKernelGGL(k_fill_const, dim3((unsigned)((h->nk_total + 255) / 256)), dim3(256), 0, h->stream, d,
                       (const int32_t *)h->klist_dev, h->cstr_dev, h->nk_total, h->cnz_dev);
    HIP_TRY(hipGetLastError());
  }
  if (ip && h->nsh > 0) {
    hipLaunchKernelGGL(k_fill_shared, dim3((unsigned)((h->nsh + 255) / 256)), dim3(256), 0, h->stream, h->hsig_dev,
                       (const double *)d.sig, d.ldsig, (const int32_t *)h->shared_dev, h->nsh);
    HIP_TRY(hipGetLastError());
  }
  f = FusedView{h->tiles_dev, h->hlist_dev, h->hid_dev, h->uloc_dev, h->cstr_dev, h->cnz_dev, h->nk_total,
                ip ? 1 : 0,   h->nsh,       nullptr,    nullptr,    nullptr,     nullptr,    h->fused_cap};
  return EVP1D_OK;
}

// After nsub in-place subcycles: the shared points' stresses (copy nsub % 2)
// back into d.sig.
static int fused_epilogue(evp1d_handle *h, const DevView &d, int nsub, bool ip) {
  if (!ip || h->nsh == 0) return EVP1D_OK;
  hipLaunchKernelGGL(k_scatter_shared, dim3((unsigned)((h->nsh + 255) / 256)), dim3(256), 0, h->stream, d.sig,
                     d.ldsig, (const double *)(h->hsig_dev + (size_t)(nsub % 2) * kShStride * h->nsh),
                     (const int32_t *)h->shared_dev, h->nsh);
  HIP_TRY(hipGetLastError());
  return EVP1D_OK;
}

// subcycle k (1-based) of an in-place run reads shared copy (k - 1) % 2 and
// writes copy k % 2
static void fused_step(evp1d_handle *h, FusedView &f, int k) {
  const size_t n = (size_t)kShStride * h->nsh;
  f.hsig_in = h->hsig_dev + ((k - 1) % 2) * n;
  f.hsig_out = h->hsig_dev + (k % 2) * n;
}

// nsub fused subcycles.  The state starts in the handle's buffers (X).
// Velocities, and stresses of a two-buffer run, are copied to the second set
// (Y) and the launches alternate so that the last one writes X: launch j of
// L writes X when L - j is even.  Stresses of an in-place run stay in X
// (the shared points' through their copy, put back at the end).  The pair
// form runs nsub / 2 two-subcycle launches, after one one-subcycle launch
// when nsub is odd; its stresses ping-pong in pairs between sig_pair and
// sig_alt (packed before its first launch, unpacked into X after the last).
// traj (pair form only; the gradient's forward sweep): also store the stepu
// slots' velocities after subcycles 1 .. nsub - 1 at traj + k * nu, k the
// subcycle, nu = u_end - u_begin (the state after the last stays in X only).
// The pair launches j0..L of a run as one cooperative launch (k_pair_persistent):
// as many workgroups as are co-resident (a multiple of the XCD count), at
// most one per tile.
static int launch_pair_persistent(evp1d_handle *h, const DevView &x, PairView pv, double2 *sg0, double2 *sg1,
                                  int j0, int L, size_t lds) {
  const int b = pair_block_index(h->pair_block);
  const PairPersistFn fn = kPairPersistFns[b][h->tile_xcd ? 1 : 0][h->v.nhalo > 0 ? 1 : 0];
  int per_cu = 0, cus = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, h->pair_block, lds));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
  if (per_cu <= 0 || cus <= 0) return fail(EVP1D_EHIP, "no co-resident workgroups for k_pair_persistent");
  int grid = std::min(per_cu * cus, h->fused_ntiles);
  if (grid >= kXcds) grid -= grid % kXcds;
  const size_t cnt_bytes = (size_t)2 * kXcds * kCntPad * sizeof(int);
  if (!h->pcnt_dev) HIP_TRY(hipMalloc((void **)&h->pcnt_dev, cnt_bytes));
  HIP_TRY(hipMemsetAsync(h->pcnt_dev, 0, cnt_bytes, h->stream));
  DevView d = x;
  PersistArgs a{h->vel_alt, sg0, sg1, h->pcnt_dev, j0, L, h->fused_ntiles};
  void *args[] = {&d, &pv, &a};
  HIP_TRY(hipLaunchCooperativeKernel((const void *)fn, dim3(grid), dim3(h->pair_block), args, lds, h->stream));
  return EVP1D_OK;
}

// persistent: the pair launches as one cooperative launch (eager only: a
// cooperative launch is not captured into the run's graph)
// sig_src (pair form, even nsub): the starting stresses are read from there
// (a stored state of a checkpointed gradient) instead of the handle's own;
// unpack = false leaves the handle's stresses stale after the run (the
// gradient's segment sweeps, whose end state nothing reads).  Within a
// gradient call (pconst_fresh) the constant entries' str1..8 and *cnz are
// filled by the first run only: they do not change during the call.
static int issue_fused(evp1d_handle *h, int nsub, double2 *traj = nullptr, bool persistent = false,
                       const double *sig_src = nullptr, bool unpack = true) {
  const bool pair = pair_form(h);
  if (traj && !pair) return fail(EVP1D_ESTATE, "issue_fused: a trajectory needs the pair form");
  if (persistent && (!pair || traj)) return fail(EVP1D_ESTATE, "issue_fused: the persistent form needs the pair form");
  const bool odd_first = pair && nsub % 2 == 1;
  if ((sig_src || !unpack) && (!pair || odd_first))
    return fail(EVP1D_ESTATE, "issue_fused: stresses in a stored state need the pair form and an even run");
  const bool fills = !(h->pconst_fresh && h->fills_fresh);
  const long nu = (long)h->v.u_end - h->v.u_begin;
  const bool ip = h->fused_inplace && !pair;
  const DevView X = h->v;
  DevView Y = h->v;
  Y.vel = Y.vel_out = h->vel_alt;
  Y.velc = h->vel_alt;
  HIP_TRY(hipMemcpyAsync(Y.vel, X.vel, (size_t)h->navel * sizeof(double2), hipMemcpyDeviceToDevice, h->stream));
  if (!ip && (!pair || odd_first)) {
    Y.sig = h->sig_alt;
    HIP_TRY(hipMemcpyAsync(Y.sig, X.sig, (size_t)12 * h->ldsig * sizeof(double), hipMemcpyDeviceToDevice,
                           h->stream));
  }
  FusedView f;
  RC_TRY(fused_prologue(h, X, f, ip, fills));
  const long t_end = h->v.t_end;
  PairView pv{h->ptiles_dev, h->hlist_dev, h->pxlist_dev, h->pelist_dev, h->puloc_dev, h->pvloc_dev,
              h->pcstr_dev,  h->cnz_dev,   h->pk_total,   nullptr,       h->pair_cap,  h->pconst,
              (long)h->na,   nullptr,      nullptr,       h->ldsig,      h->ptnbr_dev,  h->pfix_dev};
  if (pair && h->na > 0 && !h->pconst_fresh) {
    hipLaunchKernelGGL(k_pack_const, dim3(blocks_for(h->na)), dim3(kBlock), 0, h->stream, X, h->pconst, (long)h->na);
    HIP_TRY(hipGetLastError());
  }
  double2 *P[2] = {reinterpret_cast<double2 *>(h->sig_pair), reinterpret_cast<double2 *>(h->sig_alt)};
  int pin = 0;                 // P[pin]: the pair launches' current stresses
  bool packed = false;
  const double *soa = sig_src ? sig_src : X.sig;   // the current stresses before packing
  if (pair && h->pk_total > 0 && fills) {
    hipLaunchKernelGGL(k_fill_const, dim3((unsigned)((h->pk_total + 255) / 256)), dim3(256), 0, h->stream, X,
                       (const int32_t *)h->pklist_dev, h->pcstr_dev, h->pk_total, h->cnz_dev);
    HIP_TRY(hipGetLastError());
  }
  const size_t lds = (size_t)4 * h->fused_cap * sizeof(double2);
  const size_t lds2 = pair_lds(h->pair_cap, h->pair_capv);
  const int L = pair ? nsub / 2 + nsub % 2 : nsub;
  for (int j = 1; j <= L; ++j) {
    const bool to_x = ((L - j) % 2) == 0;
    DevView in = to_x ? Y : X;
    DevView out = to_x ? X : Y;
    if (pair && (j > 1 || !odd_first)) {
      if (!packed) {
        if (t_end > 0) {
          hipLaunchKernelGGL(k_pack_sig, dim3(blocks_for((int)t_end)), dim3(kBlock), 0, h->stream, P[0], h->ldsig,
                             soa, h->ldsig, t_end);
          HIP_TRY(hipGetLastError());
        }
        packed = true;
      }
      if (persistent) {   // this and every later launch, in one
        if (h->fused_ntiles > 0) RC_TRY(launch_pair_persistent(h, X, pv, P[0], P[1], j, L, lds2));
        pin = (pin + (L - j + 1)) % 2;
        break;
      }
      pv.vel_out = out.vel;
      pv.sg_in = P[pin];
      pv.sg_out = P[1 - pin];
      pin = 1 - pin;
      const int nt = pair_hints(h) ? 1 : 0;
      PairFn fn = (h->v.nhalo > 0 ? kPairFixFns : kPairFns)[pair_block_index(h->pair_block)][h->tile_xcd ? 1 : 0][nt];
      if (traj) {
        const int before = odd_first ? 2 * j - 3 : 2 * j - 2;   // subcycles done before this launch
        pv.tr_mid = traj + (size_t)(before + 1) * nu;
        pv.tr_end = before + 2 < nsub ? traj + (size_t)(before + 2) * nu : nullptr;
        fn = kPairTrajFns[pair_block_index(h->pair_block)][nt];
      }
      if (h->fused_ntiles > 0) {
        hipLaunchKernelGGL(fn, dim3(h->fused_ntiles), dim3(h->pair_block), lds2, h->stream, in, pv, j == L ? 1 : 0);
        HIP_TRY(hipGetLastError());
      }
      if (h->v.nhalo > 0) {   // the copy after the launch's second subcycle
        hipLaunchKernelGGL(k_halo, dim3(blocks_for(h->v.nhalo)), dim3(kBlock), 0, h->stream, out);
        HIP_TRY(hipGetLastError());
      }
      continue;
    }
    f.vel_out = out.vel;
    f.sig_out = out.sig;
    soa = out.sig;
    fused_step(h, f, j);
    if (h->fused_ntiles > 0) {
      hipLaunchKernelGGL(kFusedFns[fused_block_index(h->tile_block)][h->tile_xcd ? 1 : 0], dim3(h->fused_ntiles),
                         dim3(h->tile_block), lds, h->stream, in, f, j == L ? 1 : 0);
      HIP_TRY(hipGetLastError());
    }
    if (h->v.nhalo > 0) {
      hipLaunchKernelGGL(k_halo, dim3(blocks_for(h->v.nhalo)), dim3(kBlock), 0, h->stream, out);
      HIP_TRY(hipGetLastError());
    }
    if (traj && nsub > 1 && nu > 0)   // the odd first subcycle of a pair-form sweep
      HIP_TRY(hipMemcpyAsync(traj + nu, out.vel + h->v.u_begin, (size_t)nu * sizeof(double2),
                             hipMemcpyDeviceToDevice, h->stream));
  }
  if (packed && t_end > 0 && unpack) {
    hipLaunchKernelGGL(k_unpack_sig, dim3(blocks_for((int)t_end)), dim3(kBlock), 0, h->stream, X.sig, h->ldsig,
                       (const double2 *)P[pin], h->ldsig, t_end);
    HIP_TRY(hipGetLastError());
  }
  if (h->pconst_fresh) h->fills_fresh = true;
  return fused_epilogue(h, X, nsub, ip);
}

static int run_graph(evp1d_handle *h, int nsub, bool fused) {
  const int gmode = fused ? EVP1D_MODE_FUSED : EVP1D_MODE_GRAPH;
  if (h->graph_dirty || h->graph_nsub != nsub || h->graph_mode != gmode || !h->graph_exec) {
    drop_graph(h);
    if (fused) RC_TRY(ensure_fused_buffers(h));
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = fused ? issue_fused(h, nsub) : issue_subcycles(h, nsub);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(EVP1D_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    h->graph = g;
    HIP_TRY(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
    h->graph_nsub = nsub;
    h->graph_mode = gmode;
    h->graph_dirty = false;
  }
  HIP_TRY(hipGraphLaunch(h->graph_exec, h->stream));
  return EVP1D_OK;
}

static int run_persistent(evp1d_handle *h, int nsub) {
  if (h->fused_ok && pair_form(h)) {   // the pair form's tiles, one grid barrier per two subcycles
    RC_TRY(ensure_fused_buffers(h));
    return issue_fused(h, nsub, nullptr, true);
  }
  if (h->coop_blocks <= 0) {
    int per_cu = 0, cus = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persistent, kBlock, 0));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    if (per_cu <= 0 || cus <= 0) return fail(EVP1D_EHIP, "no co-resident blocks for k_persistent");
    h->coop_blocks = per_cu * cus;
  }
  const int work = std::max(std::max(h->v.t_end, h->v.u_end - h->v.u_begin), h->v.nhalo);
  const int grid = std::max(1, std::min(h->coop_blocks, blocks_for(work)));
  DevView v = h->v;
  void *args[] = {&v, &nsub};
  HIP_TRY(hipLaunchCooperativeKernel((const void *)k_persistent, dim3(grid), dim3(kBlock), args, 0, h->stream));
  return EVP1D_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

// error channel for the host-only translation units of the library
int evp1d__set_error(int code, const char *msg) { return fail(code, "%s", msg); }

int evp1d_set_evp_parameters(int32_t ndte, double dt, evp1d_params *p) {
  (void)dt;   // unused, as in src/ice.F90:119 (dtei is commented out)
  if (!p) return fail(EVP1D_EINVAL, "evp1d_set_evp_parameters: p is NULL");
  if (ndte < 1) return fail(EVP1D_EINVAL, "evp1d_set_evp_parameters: ndte=%d < 1", ndte);
  // src/ice.F90:80-107 (parameters) and 109-130 (derived values)
  const double c1 = 1.0, c2 = 2.0;
  const double elasticDamp = 0.36, e_plasticpot = c2, e_yieldcurve = c2;
  std::memset(p, 0, sizeof(*p));
  p->ndte = ndte;
  p->epp2i = c1 / (e_plasticpot * e_plasticpot);
  p->capping = c1;
  p->e_factor = (e_yieldcurve * e_yieldcurve) / (e_plasticpot * e_plasticpot * e_plasticpot * e_plasticpot);
  p->arlx = c2 * elasticDamp * (double)ndte;
  p->arlx1i = c1 / p->arlx;
  p->brlx = (double)ndte;
  p->denom1 = c1 / (c1 + p->arlx1i);
  p->deltaminEVP = 10e-9;
  p->revp = 0.0;
  p->Ktens = 0.0;
  p->u0 = 5e-5;
  p->cosw = 1.0;
  p->sinw = 0.0;
  p->rhow = 1026.0;
  return EVP1D_OK;
}

static int create_impl(int32_t nmem, const int32_t *mna, const int32_t *mnavel, const evp1d_params *p,
                       evp1d_handle **out) {
  if (!out) return fail(EVP1D_EINVAL, "evp1d_create: out is NULL");
  *out = nullptr;
  if (!p) return fail(EVP1D_EINVAL, "evp1d_create: params is NULL");
  if (nmem < 1 || !mna || !mnavel) return fail(EVP1D_EINVAL, "evp1d_create: need nmem >= 1 and size lists");
  long na_l = 0, navel_l = 0;
  for (int k = 0; k < nmem; ++k) {
    if (mna[k] < 0 || mnavel[k] < mna[k])
      return fail(EVP1D_EINVAL, "evp1d_create: member %d needs 0 <= na <= navel (na=%d navel=%d)", k, mna[k],
                  mnavel[k]);
    na_l += mna[k];
    navel_l += mnavel[k];
  }
  if (navel_l > 0x7fffffffL) return fail(EVP1D_EINVAL, "evp1d_create: %ld slots exceed int32", navel_l);
  const int na = (int)na_l, navel = (int)navel_l;
  evp1d_handle *h = new (std::nothrow) evp1d_handle();
  if (!h) return fail(EVP1D_ENOMEM, "evp1d_create: host allocation failed");
  h->na = na;
  h->navel = navel;
  {
    int a = 0, g = na;
    for (int k = 0; k < nmem; ++k) {
      h->m_na.push_back(mna[k]);
      h->m_navel.push_back(mnavel[k]);
      h->a_off.push_back(a);
      h->g_off.push_back(g);
      a += mna[k];
      g += mnavel[k] - mna[k];
    }
    for (int w = 0; w < 6; ++w) h->idx_mset[w].assign((size_t)nmem, 0);
  }
  h->p = *p;
  h->ldsig = ((long)na + 63) / 64 * 64;
  int rc = EVP1D_OK;
  auto chk = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && rc == EVP1D_OK) rc = fail(EVP1D_EHIP, "%s: %s", what, hipGetErrorString(e));
  };
  chk(hipGetDevice(&h->device), "hipGetDevice");
  // this device's address of the fixed scratch of evp1d_loss_noalloc (looked
  // up once per device, the device being current); evp1d_loss_noalloc
  // retries and reports a failure
  if (!rc) (void)scratch_setup(h->device, true);
  chk(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
  chk(hipEventCreate(&h->ev0), "hipEventCreate");
  chk(hipEventCreate(&h->ev1), "hipEventCreate");
  if (rc) { release(h); return rc; }

  DevView &v = h->v;
  v.na = na;
  v.navel = navel;
  v.ldsig = h->ldsig;
  const size_t nA = (size_t)na * sizeof(double), nV = (size_t)navel * sizeof(double2);
  auto A = [&](void **ptr, size_t bytes) { if (!rc) rc = dev_alloc(h, ptr, bytes); };
  A((void **)&v.vel, nV);
  A((void **)&v.s15, nV); A((void **)&v.s27, nV); A((void **)&v.s36, nV); A((void **)&v.s48, nV);
  A((void **)&v.sig, (size_t)12 * h->ldsig * sizeof(double));
  const double **cplain[] = {&v.dxT, &v.dyT, &v.strength, &v.hte, &v.htn, &v.htem1, &v.htnm1,
                             &v.cw, &v.aiu, &v.uocn, &v.vocn, &v.waterx, &v.watery, &v.forcex,
                             &v.forcey, &v.umassdti, &v.fm, &v.uarear, &v.uvel_init, &v.vvel_init,
                             &v.tbu};
  for (const double **pp : cplain) A((void **)pp, nA);
  A((void **)&v.strintx, nA); A((void **)&v.strinty, nA); A((void **)&v.cb, nA);
  for (int k = 0; k < 6; ++k) A((void **)&h->idx_dev[k], (size_t)na * sizeof(int32_t));
  A((void **)&h->inv_dev, (size_t)na * sizeof(int32_t));
  A((void **)&h->perm_dev, (size_t)na * sizeof(int32_t));
  for (int k = 0; k < 6; ++k) A((void **)&h->idx_caller_dev[k], (size_t)na * sizeof(int32_t));
  A((void **)&h->map_dev, (size_t)na * sizeof(int32_t));
  A((void **)&h->scratch, (size_t)navel * sizeof(double));
  A((void **)&h->ptmp, (size_t)na * sizeof(double2));
  if (rc) { release(h); return rc; }
  v.vel_out = v.vel;
  v.velc = v.vel;
  v.s15c = v.s15; v.s27c = v.s27; v.s36c = v.s36; v.s48c = v.s48;
  v.ee = h->idx_dev[0]; v.ne = h->idx_dev[1]; v.se = h->idx_dev[2];
  v.nw = h->idx_dev[3]; v.sw = h->idx_dev[4]; v.sse = h->idx_dev[5];
  v.nhalo = 0;
  v.s = scalars_from(*p);
  // no skips yet: every point is in the "stress and stepu" class, identity order
  v.t_end = na;
  v.u_begin = 0;
  v.u_end = na;
  h->skip_host.assign((size_t)na, 0);
  h->perm.resize((size_t)na);
  h->inv.resize((size_t)na);
  for (int q = 0; q < na; ++q) h->perm[(size_t)q] = h->inv[(size_t)q] = q;
  if (na > 0 &&
      hipMemcpyAsync(h->inv_dev, h->inv.data(), (size_t)na * sizeof(int32_t), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
    release(h);
    return fail(EVP1D_EHIP, "evp1d_create: upload failed");
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    release(h);
    return fail(EVP1D_EHIP, "evp1d_create: initial clear failed");
  }
  *out = h;
  return EVP1D_OK;
}

int evp1d_create(int32_t na, int32_t navel, const evp1d_params *p, evp1d_handle **out) {
  return create_impl(1, &na, &navel, p, out);
}

int evp1d_create_batch(int32_t nmem, const int32_t *na, const int32_t *navel, const evp1d_params *p,
                       evp1d_handle **out) {
  return create_impl(nmem, na, navel, p, out);
}

int evp1d_member_info(const evp1d_handle *h, int32_t k, int32_t *na, int32_t *navel, int32_t *active_offset,
                      int32_t *ghost_offset) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_member_info: NULL handle");
  MemberSpan m;
  if (k < 0 || !member_span(h, k, m)) return fail(EVP1D_EINVAL, "evp1d_member_info: no member %d", k);
  if (na) *na = m.m_na;
  if (navel) *navel = m.m_navel;
  if (active_offset) *active_offset = m.a_off;
  if (ghost_offset) *ghost_offset = m.g_off;
  return EVP1D_OK;
}

int evp1d_get_members(const evp1d_handle *h, int32_t *nmem) {
  if (!h || !nmem) return fail(EVP1D_EINVAL, "evp1d_get_members: NULL argument");
  *nmem = (int32_t)h->m_na.size();
  return EVP1D_OK;
}

int evp1d_destroy(evp1d_handle *h) {
  release(h);
  return EVP1D_OK;
}

int evp1d_set_params(evp1d_handle *h, const evp1d_params *p) {
  if (!h || !p) return fail(EVP1D_EINVAL, "evp1d_set_params: NULL argument");
  // the captured graph holds the scalars: capture again only if one changed
  // (CICE hands the same parameters to every time step)
  evp1d_params a = h->p, b = *p;
  a.reserved = b.reserved = 0;
  if (std::memcmp(&a, &b, sizeof(evp1d_params)) != 0) h->graph_dirty = true;
  h->p = *p;
  h->v.s = scalars_from(*p);
  return EVP1D_OK;
}

int evp1d_get_params(const evp1d_handle *h, evp1d_params *p) {
  if (!h || !p) return fail(EVP1D_EINVAL, "evp1d_get_params: NULL argument");
  *p = h->p;
  return EVP1D_OK;
}

int evp1d_get_sizes(const evp1d_handle *h, int32_t *na, int32_t *navel) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_sizes: NULL handle");
  if (na) *na = h->na;
  if (navel) *navel = h->navel;
  return EVP1D_OK;
}

static int set_field_impl(evp1d_handle *h, int k, int32_t field, const double *src, int32_t n) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_field: NULL handle");
  FieldRef r;
  MemberSpan m;
  if (!field_ref(h, field, r)) return fail(EVP1D_EINVAL, "evp1d_set_field: unknown field %d", field);
  if (!member_span(h, k, m)) return fail(EVP1D_EINVAL, "evp1d_set_field: no member %d", k);
  const int len = r.len == h->navel && r.kind == FK_PAIR ? m.m_navel : m.m_na;
  if (n != len) return fail(EVP1D_EINVAL, "evp1d_set_field: field %d has length %d, got %d", field, len, n);
  if (n == 0) return EVP1D_OK;
  if (!src) return fail(EVP1D_EINVAL, "evp1d_set_field: src is NULL");
  // the caller->slot map always matches the data, so no relayout is needed
  
  const Span sp{m.a_off, m.m_na, m.g_off, n};
  HIP_TRY(hipMemcpyAsync(h->scratch, src, (size_t)n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  if (r.kind == FK_PLAIN)
    hipLaunchKernelGGL(k_scatter_plain, dim3(blocks_for(n)), dim3(kBlock), 0, h->stream, r.plain,
                       (const double *)h->scratch, (const int32_t *)h->inv_dev, h->na, sp);
  else
    hipLaunchKernelGGL(k_scatter_comp, dim3(blocks_for(n)), dim3(kBlock), 0, h->stream, r.pair,
                       (const double *)h->scratch, (const int32_t *)h->inv_dev, h->na, sp, r.comp);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

static int get_field_impl(evp1d_handle *h, int k, int32_t field, double *dst, int32_t n) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_field: NULL handle");
  FieldRef r;
  MemberSpan m;
  if (!field_ref(h, field, r)) return fail(EVP1D_EINVAL, "evp1d_get_field: unknown field %d", field);
  if (!member_span(h, k, m)) return fail(EVP1D_EINVAL, "evp1d_get_field: no member %d", k);
  const int len = r.len == h->navel && r.kind == FK_PAIR ? m.m_navel : m.m_na;
  if (n != len) return fail(EVP1D_EINVAL, "evp1d_get_field: field %d has length %d, got %d", field, len, n);
  if (n == 0) return EVP1D_OK;
  if (!dst) return fail(EVP1D_EINVAL, "evp1d_get_field: dst is NULL");
  // the caller->slot map always matches the data, so no relayout is needed
  
  const Span sp{m.a_off, m.m_na, m.g_off, n};
  if (r.kind == FK_PLAIN)
    hipLaunchKernelGGL(k_gather_plain, dim3(blocks_for(n)), dim3(kBlock), 0, h->stream, h->scratch,
                       (const double *)r.plain, (const int32_t *)h->inv_dev, h->na, sp);
  else
    hipLaunchKernelGGL(k_gather_comp, dim3(blocks_for(n)), dim3(kBlock), 0, h->stream, h->scratch,
                       (const double2 *)r.pair, (const int32_t *)h->inv_dev, h->na, sp, r.comp);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(dst, h->scratch, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

static int set_index_impl(evp1d_handle *h, int k, int32_t which, const int32_t *idx, int32_t n) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_index: NULL handle");
  MemberSpan m;
  if (!member_span(h, k, m)) return fail(EVP1D_EINVAL, "evp1d_set_index: no member %d", k);
  if (which < EVP1D_EE || which > EVP1D_SSE) return fail(EVP1D_EINVAL, "evp1d_set_index: unknown list %d", which);
  if (n != m.m_na) return fail(EVP1D_EINVAL, "evp1d_set_index: length %d != na %d", n, m.m_na);
  if (n > 0 && !idx) return fail(EVP1D_EINVAL, "evp1d_set_index: idx is NULL");
  std::vector<int32_t> &dst = h->idx_host[which];
  if (dst.size() != (size_t)h->na) dst.assign((size_t)h->na, 0);
  for (int i = 0; i < n; ++i) {
    const int32_t q = idx[i];
    if (q < 1 || q > m.m_navel)
      return fail(EVP1D_EINDEX, "evp1d_set_index: list %d entry %d = %d outside 1..%d", which, i + 1, q, m.m_navel);
  }
  for (int i = 0; i < n; ++i) {
    const int32_t q = idx[i];   // member-local, 1-based
    dst[(size_t)m.a_off + i] = q <= m.m_na ? m.a_off + q - 1 : m.g_off + (q - m.m_na - 1);
  }
  std::vector<uint8_t> &ms = h->idx_mset[which];
  if (k < 0) std::fill(ms.begin(), ms.end(), (uint8_t)1); else ms[(size_t)k] = 1;
  h->idx_set[which] = std::all_of(ms.begin(), ms.end(), [](uint8_t b) { return b != 0; });
  h->idx_dirty = true;
  h->layout_dirty = true;   // the tiled slot order follows the lists
  h->tile_cache_ok = false;
  h->coords_ok = false;
  h->idx_caller_dirty[which] = true;
  return EVP1D_OK;
}

static int set_skip_impl(evp1d_handle *h, int k, int32_t which, const int32_t *flags, int32_t n) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_skip: NULL handle");
  MemberSpan m;
  if (!member_span(h, k, m)) return fail(EVP1D_EINVAL, "evp1d_set_skip: no member %d", k);
  if (which != EVP1D_SKIPT && which != EVP1D_SKIPU) return fail(EVP1D_EINVAL, "evp1d_set_skip: unknown flag %d", which);
  if (n != m.m_na) return fail(EVP1D_EINVAL, "evp1d_set_skip: length %d != na %d", n, m.m_na);
  if (n > 0 && !flags) return fail(EVP1D_EINVAL, "evp1d_set_skip: flags is NULL");
  const uint8_t bit = (uint8_t)(1u << which);
  bool changed = false;
  for (int i = 0; i < n; ++i) {
    uint8_t &b = h->skip_host[(size_t)m.a_off + i];
    const uint8_t nb = (uint8_t)(flags[i] ? (b | bit) : (b & ~bit));
    changed |= nb != b;
    b = nb;
  }
  if (changed) h->layout_dirty = true;
  return EVP1D_OK;
}

int evp1d_set_field(evp1d_handle *h, int32_t field, const double *src, int32_t n) {
  return set_field_impl(h, -1, field, src, n);
}
int evp1d_get_field(evp1d_handle *h, int32_t field, double *dst, int32_t n) {
  return get_field_impl(h, -1, field, dst, n);
}
int evp1d_set_index(evp1d_handle *h, int32_t which, const int32_t *idx, int32_t n) {
  return set_index_impl(h, -1, which, idx, n);
}
int evp1d_set_skip(evp1d_handle *h, int32_t which, const int32_t *flags, int32_t n) {
  return set_skip_impl(h, -1, which, flags, n);
}
int evp1d_set_field_member(evp1d_handle *h, int32_t k, int32_t field, const double *src, int32_t n) {
  if (k < 0) return fail(EVP1D_EINVAL, "evp1d_set_field_member: member %d < 0", k);
  return set_field_impl(h, k, field, src, n);
}
int evp1d_get_field_member(evp1d_handle *h, int32_t k, int32_t field, double *dst, int32_t n) {
  if (k < 0) return fail(EVP1D_EINVAL, "evp1d_get_field_member: member %d < 0", k);
  return get_field_impl(h, k, field, dst, n);
}
int evp1d_set_index_member(evp1d_handle *h, int32_t k, int32_t which, const int32_t *idx, int32_t n) {
  if (k < 0) return fail(EVP1D_EINVAL, "evp1d_set_index_member: member %d < 0", k);
  return set_index_impl(h, k, which, idx, n);
}
int evp1d_set_skip_member(evp1d_handle *h, int32_t k, int32_t which, const int32_t *flags, int32_t n) {
  if (k < 0) return fail(EVP1D_EINVAL, "evp1d_set_skip_member: member %d < 0", k);
  return set_skip_impl(h, k, which, flags, n);
}

int evp1d_set_halo(evp1d_handle *h, int32_t n, const int32_t *dst, const int32_t *src) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_halo: NULL handle");
  if (n < 0) return fail(EVP1D_EINVAL, "evp1d_set_halo: n=%d < 0", n);
  if (n > 0 && (!dst || !src)) return fail(EVP1D_EINVAL, "evp1d_set_halo: NULL list");
  std::vector<int32_t> d((size_t)n), s((size_t)n);
  std::vector<char> seen((size_t)(h->navel - h->na), 0);
  for (int k = 0; k < n; ++k) {
    if (dst[k] <= h->na || dst[k] > h->navel)
      return fail(EVP1D_EINDEX, "evp1d_set_halo: dst[%d]=%d not a ghost slot (%d..%d)", k + 1, dst[k], h->na + 1,
                  h->navel);
    if (src[k] < 0 || src[k] > h->na)
      return fail(EVP1D_EINDEX, "evp1d_set_halo: src[%d]=%d outside 0..%d", k + 1, src[k], h->na);
    char &seen_k = seen[(size_t)(dst[k] - h->na - 1)];
    if (seen_k) return fail(EVP1D_EINDEX, "evp1d_set_halo: ghost slot %d listed twice", dst[k]);
    seen_k = 1;
    d[(size_t)k] = dst[k] - 1;
    s[(size_t)k] = src[k] - 1;   // -1: fill with zero
  }
  if (n > h->halo_cap) {
    RC_TRY(dev_alloc(h, (void **)&h->halo_dev, (size_t)2 * n * sizeof(int32_t)));
    h->halo_cap = n;
  }
  h->halo_dst_host.swap(d);
  h->halo_src_host.swap(s);
  h->v.nhalo = n;
  h->v.halo_dst = h->halo_dev;
  h->v.halo_src = h->halo_dev ? h->halo_dev + n : nullptr;
  h->idx_dirty = true;
  h->layout_dirty = true;   // the pair form's level-2 reads of listed ghosts change
  return EVP1D_OK;
}

int evp1d_stress(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_stress: NULL handle");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_stress: neighbour lists not set");
  RC_TRY(ensure(h));
  return launch_stress(h);
}

int evp1d_stepu(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_stepu: NULL handle");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_stepu: neighbour lists not set");
  RC_TRY(ensure(h));
  return launch_stepu(h, true);
}

int evp1d_halo(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_halo: NULL handle");
  RC_TRY(ensure(h));
  return launch_halo(h);
}

int evp1d_set_mode(evp1d_handle *h, int32_t mode) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_mode: NULL handle");
  if (mode < EVP1D_MODE_LAUNCH || mode > EVP1D_MODE_FUSED) return fail(EVP1D_EINVAL, "evp1d_set_mode: bad mode %d", mode);
  h->mode = mode;
  return EVP1D_OK;
}

int evp1d_get_mode(const evp1d_handle *h, int32_t *mode) {
  if (!h || !mode) return fail(EVP1D_EINVAL, "evp1d_get_mode: NULL argument");
  *mode = h->mode;
  return EVP1D_OK;
}

int evp1d_run(evp1d_handle *h, int32_t nsub) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_run: NULL handle");
  if (nsub < 0) return fail(EVP1D_EINVAL, "evp1d_run: nsub=%d < 0", nsub);
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_run: neighbour lists not set");
  RC_TRY(ensure(h));
  if (nsub == 0 || h->na == 0) return EVP1D_OK;
  switch (h->mode) {
    case EVP1D_MODE_LAUNCH: return issue_subcycles(h, nsub);
    case EVP1D_MODE_PERSISTENT: return run_persistent(h, nsub);
    case EVP1D_MODE_FUSED:
      if (!h->fused_ok) return fail(EVP1D_ESTATE, "evp1d_run: fused form unavailable (%s)", h->fused_why.c_str());
      return run_graph(h, nsub, true);
    case EVP1D_MODE_GRAPH: return run_graph(h, nsub, false);
    case EVP1D_MODE_AUTO:
    default:
      // the fused form copies the state once per run (two buffers), which
      // only pays off over several subcycles
      return run_graph(h, nsub, h->fused_ok && nsub >= kFusedMinSub);
  }
}

int evp1d_run_individual(evp1d_handle *h, int32_t which, int32_t nsub) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_run_individual: NULL handle");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_run_individual: neighbour lists not set");
  if (which < 0 || which > 2) return fail(EVP1D_EINVAL, "evp1d_run_individual: which=%d", which);
  RC_TRY(ensure(h));
  if (which == 2) {
    if (!h->fused_ok)
      return fail(EVP1D_ESTATE, "evp1d_run_individual: fused form unavailable (%s)", h->fused_why.c_str());
    if (nsub <= 0) return EVP1D_OK;
    RC_TRY(ensure_fused_buffers(h));
    return issue_fused(h, nsub);
  }
  for (int k = 0; k < nsub; ++k) RC_TRY(which == 0 ? launch_stress(h) : launch_stepu(h, k == nsub - 1));
  return EVP1D_OK;
}

static int block_index(int32_t block) {
  for (int i = 0; i < kNumBlocks; ++i)
    if (kBlocks[i] == block) return i;
  return -1;
}

int evp1d_set_tuning(evp1d_handle *h, int32_t stress_block, int32_t stress_xcd, int32_t stepu_block,
                     int32_t stepu_xcd) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_tuning: NULL handle");
  const int a = block_index(stress_block), b = block_index(stepu_block);
  if (a < 0 || b < 0) return fail(EVP1D_EINVAL, "evp1d_set_tuning: block sizes must be 64, 128, 256, 512 or 1024");
  h->cfg_stress = LaunchCfg{a, stress_xcd ? 1 : 0};
  h->cfg_stepu = LaunchCfg{b, stepu_xcd ? 1 : 0};
  h->graph_dirty = true;
  return EVP1D_OK;
}

int evp1d_get_tuning(const evp1d_handle *h, int32_t *stress_block, int32_t *stress_xcd, int32_t *stepu_block,
                     int32_t *stepu_xcd) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_tuning: NULL handle");
  if (stress_block) *stress_block = kBlocks[h->cfg_stress.bi];
  if (stress_xcd) *stress_xcd = h->cfg_stress.xcd;
  if (stepu_block) *stepu_block = kBlocks[h->cfg_stepu.bi];
  if (stepu_xcd) *stepu_xcd = h->cfg_stepu.xcd;
  return EVP1D_OK;
}

// Time every launch configuration of k_stress and k_stepu on this handle's
// own data (reps launches each, HIP events) and keep the fastest.  The state
// the trial launches change is saved first and restored afterwards, so the
// handle's results are unaffected.
int evp1d_set_tiling(evp1d_handle *h, int32_t tile_w, int32_t tile_h, int32_t block, int32_t xcd) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_tiling: NULL handle");
  if (tile_w < 1 || tile_h < 1 || (long)(tile_w + 1) * (tile_h + 1) > kFusedMaxEntries)
    return fail(EVP1D_EINVAL, "evp1d_set_tiling: tile %d x %d: need w, h >= 1 and (w+1)(h+1) <= %d", tile_w, tile_h,
                kFusedMaxEntries);
  if (xcd != 0 && xcd != 1) return fail(EVP1D_EINVAL, "evp1d_set_tiling: xcd must be 0 or 1");
  if (fused_block_index(block) < 0) return fail(EVP1D_EINVAL, "evp1d_set_tiling: block must be 64, 128, 256 or 512");
  h->tile_block = block;
  if (tile_w != h->tile_w || tile_h != h->tile_h || h->tiling_auto) {
    h->layout_dirty = true;
    h->tile_cache_ok = false;
  }
  h->tiling_auto = false;
  h->tile_w = tile_w;
  h->tile_h = tile_h;
  h->tile_xcd = xcd;
  h->graph_dirty = true;
  return EVP1D_OK;
}

int evp1d_set_tile_order(evp1d_handle *h, int32_t order) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_tile_order: NULL handle");
  if (order != 0 && order != 1) return fail(EVP1D_EINVAL, "evp1d_set_tile_order: order must be 0 or 1");
  if (order != h->tile_order) {
    h->tile_order = order;
    h->layout_dirty = true;
    h->tile_cache_ok = false;
    h->graph_dirty = true;
  }
  return EVP1D_OK;
}

int evp1d_get_tile_order(evp1d_handle *h, int32_t *order) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_tile_order: NULL handle");
  if (order) *order = h->tile_order;
  return EVP1D_OK;
}

int evp1d_get_tiling(evp1d_handle *h, int32_t *tile_w, int32_t *tile_h, int32_t *block, int32_t *xcd,
                     int32_t *ntiles) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_tiling: NULL handle");
  if (ready(h)) RC_TRY(ensure(h));
  if (tile_w) *tile_w = h->tile_w;
  if (tile_h) *tile_h = h->tile_h;
  if (block) *block = h->tile_block;
  if (xcd) *xcd = h->tile_xcd;
  if (ntiles) *ntiles = h->fused_ok ? h->fused_ntiles : 0;
  return EVP1D_OK;
}

extern "C++" {
// Device copies of the evolving state (stresses, velocities, str1..8,
// diagnostics) in the current slot order, for timing runs that must leave
// the state as it was.
struct SavedState {
  struct Buf { void *dev; void *bak; size_t bytes; };
  std::vector<Buf> bufs;
  int save(evp1d_handle *h, const char *who) {
    auto one = [&](void *p, size_t bytes) -> int {
      if (bytes == 0) return EVP1D_OK;
      void *b = nullptr;
      const hipError_t e = hipMalloc(&b, bytes);
      if (e != hipSuccess) return fail(EVP1D_ENOMEM, "%s: hipMalloc(%zu): %s", who, bytes, hipGetErrorString(e));
      bufs.push_back(Buf{p, b, bytes});
      HIP_TRY(hipMemcpyAsync(b, p, bytes, hipMemcpyDeviceToDevice, h->stream));
      return EVP1D_OK;
    };
    const size_t nV = (size_t)h->navel * sizeof(double2), nA = (size_t)h->na * sizeof(double);
    RC_TRY(one(h->v.sig, (size_t)12 * h->ldsig * sizeof(double)));
    void *pairs[5] = {h->v.vel, h->v.s15, h->v.s27, h->v.s36, h->v.s48};
    for (void *q : pairs) RC_TRY(one(q, nV));
    void *diag[3] = {h->v.strintx, h->v.strinty, h->v.cb};
    for (void *q : diag) RC_TRY(one(q, nA));
    return EVP1D_OK;
  }
  // copies back (keeping rc's first error), waits, frees the copies
  int restore(evp1d_handle *h, int rc, const char *who) {
    for (const Buf &b : bufs)
      if (hipMemcpyAsync(b.dev, b.bak, b.bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess && !rc)
        rc = fail(EVP1D_EHIP, "%s: restore failed", who);
    if (hipStreamSynchronize(h->stream) != hipSuccess && !rc) rc = fail(EVP1D_EHIP, "%s: sync failed", who);
    for (const Buf &b : bufs) (void)hipFree(b.bak);
    bufs.clear();
    return rc;
  }
};

// average duration in microseconds of `reps` launches of what `issue` puts
// on the handle's stream (after one untimed call)
template <class F>
static int time_launches(evp1d_handle *h, int reps, F issue, float *us) {
  RC_TRY(issue());
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  for (int r = 0; r < reps; ++r) RC_TRY(issue());
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *us = ms * 1000.f / (float)reps;
  return EVP1D_OK;
}
}  // extern "C++"

int evp1d_autotune(evp1d_handle *h, int32_t reps, float *stress_us, float *stepu_us) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_autotune: NULL handle");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_autotune: neighbour lists not set");
  if (reps < 1) return fail(EVP1D_EINVAL, "evp1d_autotune: reps=%d < 1", reps);
  RC_TRY(ensure(h));
  SavedState st;
  int rc = st.save(h, "evp1d_autotune");
  float best_s = 1e30f, best_u = 1e30f;
  const LaunchCfg keep_s = h->cfg_stress, keep_u = h->cfg_stepu;
  LaunchCfg win_s = keep_s, win_u = keep_u;
  for (int which = 0; which < 2 && !rc; ++which) {
    for (int bi = 0; bi < kNumBlocks && !rc; ++bi) {
      for (int x = 0; x < 2 && !rc; ++x) {
        (which == 0 ? h->cfg_stress : h->cfg_stepu) = LaunchCfg{bi, x};
        float us = 0.f;
        rc = time_launches(h, reps, [&] { return which == 0 ? launch_stress(h) : launch_stepu(h, true); }, &us);
        if (!rc && which == 0 && us < best_s) { best_s = us; win_s = h->cfg_stress; }
        if (!rc && which == 1 && us < best_u) { best_u = us; win_u = h->cfg_stepu; }
      }
    }
  }
  rc = st.restore(h, rc, "evp1d_autotune");
  h->cfg_stress = rc ? keep_s : win_s;
  h->cfg_stepu = rc ? keep_u : win_u;
  h->graph_dirty = true;
  if (!rc) {
    if (stress_us) *stress_us = best_s;
    if (stepu_us) *stepu_us = best_u;
  }
  return rc;
}

int evp1d_set_fused_inplace(evp1d_handle *h, int32_t inplace) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_fused_inplace: NULL handle");
  if (inplace != 0 && inplace != 1) return fail(EVP1D_EINVAL, "evp1d_set_fused_inplace: inplace must be 0 or 1");
  if (inplace != h->fused_inplace) {
    HIP_TRY(hipStreamSynchronize(h->stream));   // no queued run may still use the other form's buffers
    h->fused_inplace = inplace;
    h->graph_dirty = true;
  }
  return EVP1D_OK;
}

int evp1d_get_fused_inplace(evp1d_handle *h, int32_t *inplace) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_fused_inplace: NULL handle");
  if (inplace) *inplace = h->fused_inplace;
  return EVP1D_OK;
}

int evp1d_set_fused_depth(evp1d_handle *h, int32_t depth) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_fused_depth: NULL handle");
  if (depth != 1 && depth != 2) return fail(EVP1D_EINVAL, "evp1d_set_fused_depth: depth must be 1 or 2");
  if (depth != h->fused_depth) {
    HIP_TRY(hipStreamSynchronize(h->stream));   // no queued run may still use the old tables
    h->fused_depth = depth;
    h->layout_dirty = true;                     // plan (or drop) the pair tables
    h->graph_dirty = true;
    if (h->tiling_auto) h->tile_cache_ok = false;   // the library's tile depends on the depth
  }
  return EVP1D_OK;
}

int evp1d_set_adjoint_segment(evp1d_handle *h, int32_t seg) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_adjoint_segment: NULL handle");
  if (seg < 0) return fail(EVP1D_EINVAL, "evp1d_set_adjoint_segment: seg=%d < 0", seg);
  h->adj_segment = seg;
  return EVP1D_OK;
}

int evp1d_get_adjoint_segment(evp1d_handle *h, int32_t *seg) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_adjoint_segment: NULL handle");
  if (seg) *seg = h->adj_segment;
  return EVP1D_OK;
}

int evp1d_get_gradient_storage(evp1d_handle *h, int64_t *bytes) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_gradient_storage: NULL handle");
  if (bytes) *bytes = (int64_t)(h->traj_bytes + h->ckpt_bytes + h->adj_ws_bytes);
  return EVP1D_OK;
}

int evp1d_release_gradient_storage(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_release_gradient_storage: NULL handle");
  if (!h->traj && !h->ckpt && !h->adj_ws) return EVP1D_OK;
  HIP_TRY(hipStreamSynchronize(h->stream));   // no queued work may still use them
  void **bufs[3] = {&h->traj, &h->ckpt, &h->adj_ws};
  size_t *sizes[3] = {&h->traj_bytes, &h->ckpt_bytes, &h->adj_ws_bytes};
  for (int k = 0; k < 3; ++k) {
    if (*bufs[k]) HIP_TRY(hipFree(*bufs[k]));
    *bufs[k] = nullptr;
    *sizes[k] = 0;
  }
  return EVP1D_OK;
}

int evp1d_get_fused_depth(evp1d_handle *h, int32_t *depth, int32_t *active, int32_t *block) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_fused_depth: NULL handle");
  if (ready(h)) RC_TRY(ensure(h));
  if (depth) *depth = h->fused_depth;
  if (active) *active = pair_form(h) ? 1 : 0;
  if (block) *block = pair_form(h) ? h->pair_block : 0;
  return EVP1D_OK;
}

int evp1d_set_cache_hints(evp1d_handle *h, int32_t hints) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_cache_hints: NULL handle");
  if (hints < -1 || hints > 1) return fail(EVP1D_EINVAL, "evp1d_set_cache_hints: hints=%d not -1, 0 or 1", hints);
  if (hints != h->cache_hints) {
    h->cache_hints = hints;
    h->graph_dirty = true;   // the captured launches name the kernel
  }
  return EVP1D_OK;
}

int evp1d_get_cache_hints(evp1d_handle *h, int32_t *hints, int32_t *active) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_cache_hints: NULL handle");
  if (ready(h)) RC_TRY(ensure(h));
  if (hints) *hints = h->cache_hints;
  if (active) *active = pair_form(h) && pair_hints(h) ? 1 : 0;
  return EVP1D_OK;
}

int evp1d_autotune_fused(evp1d_handle *h, int32_t nsub, float *us) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_autotune_fused: NULL handle");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_autotune_fused: neighbour lists not set");
  if (nsub < 1) return fail(EVP1D_EINVAL, "evp1d_autotune_fused: nsub=%d < 1", nsub);
  RC_TRY(ensure(h));
  if (!h->fused_ok)
    return fail(EVP1D_ESTATE, "evp1d_autotune_fused: fused form unavailable (%s)", h->fused_why.c_str());
  // depth 1 with both stress storages (tiles in rows), then the pair form
  // (two buffers), mostly along the Hilbert curve.  The pair form's
  // workgroup holds the largest number of level-2 stress points of any tile
  // (256, 512 or 1024 threads), so the fastest pair tiles are the largest
  // whose busiest tile still fits a workgroup size: on the synthetic ice
  // fields 31x19, 27x23, 39x15 and 33x17 for 512 threads (a tile's (W+1) x
  // (H+1) bound exceeds 512, its ice-covered points do not), 62x19 and 62x15
  // for 1024, 15x19 and 23x11 for 256; 31x15 and 15x15 fit by their bound.
  struct Cand { int w, h, block, depth, order; };
  static const Cand cands[] = {
      {64, 16, 512, 1, 0}, {32, 16, 512, 1, 0}, {32, 8, 256, 1, 0}, {16, 8, 128, 1, 0},
      {31, 19, 512, 2, 1}, {27, 23, 512, 2, 1}, {39, 15, 512, 2, 1}, {33, 17, 512, 2, 1}, {31, 15, 512, 2, 1},
      {62, 19, 512, 2, 1}, {62, 15, 512, 2, 1}, {15, 19, 256, 2, 1}, {23, 11, 256, 2, 1}, {15, 15, 256, 2, 1},
      {62, 7, 512, 2, 1}, {31, 19, 512, 2, 0}, {31, 15, 512, 2, 0}, {15, 15, 256, 2, 0}};
  constexpr int ncand = (int)(sizeof(cands) / sizeof(cands[0]));
  const bool keep_auto = h->tiling_auto;
  const int keep_w = h->tile_w, keep_h = h->tile_h, keep_b = h->tile_block, keep_x = h->tile_xcd,
            keep_ip = h->fused_inplace, keep_d = h->fused_depth, keep_o = h->tile_order,
            keep_hints = h->cache_hints;
  float best = 1e30f;
  int win = -1, win_ip = 0, win_o = 0, win_h = -1;
  int rc = EVP1D_OK;
  for (int c = 0; c < ncand && !rc; ++c) {
    const int order = cands[c].order;
    rc = evp1d_set_fused_depth(h, cands[c].depth);
    if (!rc) rc = evp1d_set_tile_order(h, order);
    if (!rc) rc = evp1d_set_tiling(h, cands[c].w, cands[c].h, cands[c].block, 1);
    if (!rc) rc = ensure(h);
    if (rc || !h->fused_ok) continue;                       // tile too large for LDS on these lists
    if (cands[c].depth == 2 && !pair_form(h)) continue;     // pair tables do not fit
    SavedState st;
    rc = st.save(h, "evp1d_autotune_fused");
    // depth 1: both stress storages; depth 2: without and with cache hints
    for (int v = 0; v < 2 && !rc; ++v) {
      h->fused_inplace = cands[c].depth == 1 ? v : 0;
      h->cache_hints = cands[c].depth == 2 ? v : keep_hints;
      rc = ensure_fused_buffers(h);
      float t = 0.f;
      if (!rc) rc = time_launches(h, 1, [&] { return issue_fused(h, nsub); }, &t);
      t /= (float)nsub;
      if (!rc && t < best) { best = t; win = c; win_ip = h->fused_inplace; win_o = order; win_h = h->cache_hints; }
    }
    rc = st.restore(h, rc, "evp1d_autotune_fused");
  }
  h->graph_dirty = true;
  if (rc || win < 0) {
    h->cache_hints = keep_hints;
    h->fused_inplace = keep_ip;
    (void)evp1d_set_fused_depth(h, keep_d);
    (void)evp1d_set_tile_order(h, keep_o);
    if (keep_auto) {
      h->tiling_auto = true;
      h->layout_dirty = true;
      h->tile_cache_ok = false;
    } else {
      (void)evp1d_set_tiling(h, keep_w, keep_h, keep_b, keep_x);
    }
    return rc ? rc : fail(EVP1D_ESTATE, "evp1d_autotune_fused: no tiling fits these lists");
  }
  h->fused_inplace = win_ip;
  h->cache_hints = cands[win].depth == 2 ? win_h : keep_hints;
  RC_TRY(evp1d_set_fused_depth(h, cands[win].depth));
  RC_TRY(evp1d_set_tile_order(h, win_o));
  RC_TRY(evp1d_set_tiling(h, cands[win].w, cands[win].h, cands[win].block, 1));
  RC_TRY(ensure(h));
  RC_TRY(ensure_fused_buffers(h));
  if (us) *us = best;
  return EVP1D_OK;
}

// ---------------------------------------------------------------------------
// adjoint driver: binomial checkpointing over the subcycles
// ---------------------------------------------------------------------------
namespace {

// Optimal schedule for reversing n steps with s free snapshot slots when the
// state at the segment start is already stored: split at m, advance m steps,
// store, reverse the right part with s-1 slots, then the left part with s.
// cost(n, s) = min_m m + cost(n-m, s-1) + cost(m, s); cost(1, s) = 0,
// cost(n, 0) = n(n-1)/2 (restart from the stored state for every step).
struct Schedule {
  int N = 0, S = 0;
  std::vector<long> cost;
  std::vector<int> split;
  long &C(int n, int s) { return cost[(size_t)n * (S + 1) + s]; }
  int &M(int n, int s) { return split[(size_t)n * (S + 1) + s]; }
  void build(int n_max, int s_max) {
    N = n_max;
    S = s_max;
    cost.assign((size_t)(N + 1) * (S + 1), 0);
    split.assign((size_t)(N + 1) * (S + 1), 0);
    for (int n = 1; n <= N; ++n) C(n, 0) = (long)n * (n - 1) / 2;
    for (int s = 1; s <= S; ++s) {
      C(1, s) = 0;
      for (int n = 2; n <= N; ++n) {
        long best = -1;
        int bm = 1;
        for (int m = 1; m < n; ++m) {
          const long c = m + C(n - m, s - 1) + C(m, s);
          if (best < 0 || c < best) { best = c; bm = m; }
        }
        C(n, s) = best;
        M(n, s) = bm;
      }
    }
  }
};

struct StateBuf {
  double *sig = nullptr;    // [12][ldsig]
  double2 *vel = nullptr;   // [navel]
};

}  // namespace

// device bytes of a stored gradient trajectory of nsub subcycles: the
// stepu-range velocities of every subcycle (AdjointRun::run_segment)
static size_t trajectory_bytes(const evp1d_handle *h, int nsub) {
  const size_t nu = (size_t)(h->v.u_end - h->v.u_begin);
  return ((size_t)nsub * nu * sizeof(double2) + 255) / 256 * 256;
}

struct AdjointRun {
  evp1d_handle *h;
  int nsub;
  AdjView a{};
  std::vector<StateBuf> slots;
  std::vector<char> used;
  StateBuf final_state;
  double2 *str_save = nullptr;   // [4][navel]
  double *ut = nullptr, *vt = nullptr, *terms = nullptr;
  Schedule sched;
  int seg = 1;                   // subcycles per checkpointing unit (the last unit may be shorter)
  int nunits = 0;                // ceil(nsub / seg)
  double *diag_save = nullptr;   // [3][na]: strintx, strinty, Cb of x_nsub (checkpointed, seg > 1)
  int working = -1;              // unit index i of the state held by the working arrays
  // lazy: pair-form units of even length.  A restore then copies only the
  // velocities and leaves sig_at pointing at the stored state's stresses,
  // which the next sweep packs from directly; a segment sweep that is not
  // the last leaves the handle's stresses stale (working = -1).
  bool lazy = false;
  const double *sig_at = nullptr;
  double loss = 0.0;

  ~AdjointRun() {
    if (!h) return;
    h->pconst_fresh = false;
    h->fills_fresh = false;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
  }
  // The stored states, the final state and its str1..8 and diagnostics:
  // carved from one allocation that the handle keeps (large allocations and
  // frees are slow).  Every part is written before it is read: a stored
  // state by take, the final state and its saves by save_final or
  // final_step, before restore_final.
  int carve_checkpoints(int nslots) {
    const size_t nA = (size_t)h->na * sizeof(double), nV = (size_t)h->navel * sizeof(double2);
    const size_t nS = (size_t)h->na * sizeof(double2);
    auto up = [](size_t b) { return (std::max<size_t>(b, 1) + 255) / 256 * 256; };
    const size_t total = (seg > 1 ? up(3 * nA) : 0) + up(4 * nV) + up(sig_bytes()) + up(nV) +
                         (size_t)nslots * (up(sig_bytes()) + up(nS));
    if (h->ckpt_bytes < total) {
      HIP_TRY(hipStreamSynchronize(h->stream));
      if (h->ckpt) HIP_TRY(hipFree(h->ckpt));
      h->ckpt = nullptr;
      h->ckpt_bytes = 0;
      const hipError_t e = hipMalloc(&h->ckpt, total);
      if (e != hipSuccess) {
        h->ckpt = nullptr;
        return fail(EVP1D_ENOMEM, "evp1d_adjoint_run: hipMalloc(%zu) for %d stored states: %s", total, nslots,
                    hipGetErrorString(e));
      }
      h->ckpt_bytes = total;
    }
    char *p = static_cast<char *>(h->ckpt);
    auto take_bytes = [&](size_t b) {
      char *q = p;
      p += up(b);
      return q;
    };
    diag_save = seg > 1 ? reinterpret_cast<double *>(take_bytes(3 * nA)) : nullptr;
    str_save = reinterpret_cast<double2 *>(take_bytes(4 * nV));
    final_state.sig = reinterpret_cast<double *>(take_bytes(sig_bytes()));
    final_state.vel = reinterpret_cast<double2 *>(take_bytes(nV));
    slots.resize((size_t)nslots);
    used.assign((size_t)nslots, 0);
    for (StateBuf &s : slots) {
      s.sig = reinterpret_cast<double *>(take_bytes(sig_bytes()));
      s.vel = reinterpret_cast<double2 *>(take_bytes(nS));
    }
    return EVP1D_OK;
  }
  size_t sig_bytes() const { return (size_t)12 * h->ldsig * sizeof(double); }
  int copy_state(const StateBuf &dst, const StateBuf &src, bool ghosts) {
    HIP_TRY(hipMemcpyAsync(dst.sig, src.sig, sig_bytes(), hipMemcpyDeviceToDevice, h->stream));
    const size_t n = (size_t)(ghosts ? h->navel : h->na) * sizeof(double2);
    HIP_TRY(hipMemcpyAsync(dst.vel, src.vel, n, hipMemcpyDeviceToDevice, h->stream));
    return EVP1D_OK;
  }
  // the handle's arrays, and the working state's (its stresses may still be
  // in a stored state: sig_at)
  StateBuf own() const { return StateBuf{h->v.sig, h->v.vel}; }
  StateBuf work() const { return StateBuf{sig_at ? const_cast<double *>(sig_at) : h->v.sig, h->v.vel}; }
  int restore(int slot, int index) {
    if (working == index) return EVP1D_OK;
    if (lazy) {
      HIP_TRY(hipMemcpyAsync(h->v.vel, slots[(size_t)slot].vel, (size_t)h->na * sizeof(double2),
                             hipMemcpyDeviceToDevice, h->stream));
      sig_at = slots[(size_t)slot].sig;
    } else {
      RC_TRY(copy_state(own(), slots[(size_t)slot], false));
    }
    working = index;
    return EVP1D_OK;
  }
  // first subcycle of unit i (0-based; unit nunits ends at nsub)
  int unit_start(int i) const { return std::min(i * seg, nsub); }
  // n units forward on the working arrays: segments as one fused run (the
  // handle's form), single subcycles as unfused launches (no per-run set-up)
  int advance(int n) {
    const int m = unit_start(working + n) - unit_start(working);
    if (seg > 1 && m > 0 && h->fused_ok) {
      RC_TRY(ensure_fused_buffers(h));
      RC_TRY(issue_fused(h, m, nullptr, false, sig_at));
      sig_at = nullptr;
    } else {
      for (int k = 0; k < m; ++k) {
        RC_TRY(launch_stress(h));
        RC_TRY(launch_stepu(h, false));
        RC_TRY(launch_halo(h));
      }
    }
    working += n;
    return EVP1D_OK;
  }
  // x_nsub of a segmented run: its state, str1..8 and diagnostics, put back
  // into the handle after the reverse sweep
  int save_final() {
    RC_TRY(copy_state(final_state, work(), true));
    const size_t nV = (size_t)h->navel * sizeof(double2), nA = (size_t)h->na * sizeof(double);
    double2 *pairs[4] = {h->v.s15, h->v.s27, h->v.s36, h->v.s48};
    for (int k = 0; k < 4; ++k)
      HIP_TRY(hipMemcpyAsync(str_save + (size_t)k * h->navel, pairs[k], nV, hipMemcpyDeviceToDevice, h->stream));
    double *dg[3] = {h->v.strintx, h->v.strinty, h->v.cb};
    for (int k = 0; k < 3; ++k)
      HIP_TRY(hipMemcpyAsync(diag_save + (size_t)k * h->na, dg[k], nA, hipMemcpyDeviceToDevice, h->stream));
    return EVP1D_OK;
  }
  int restore_final() {
    RC_TRY(copy_state(own(), final_state, false));
    const size_t nV = (size_t)h->navel * sizeof(double2), nA = (size_t)h->na * sizeof(double);
    double2 *pairs[4] = {h->v.s15, h->v.s27, h->v.s36, h->v.s48};
    for (int k = 0; k < 4; ++k)
      HIP_TRY(hipMemcpyAsync(pairs[k], str_save + (size_t)k * h->navel, nV, hipMemcpyDeviceToDevice, h->stream));
    if (diag_save) {
      double *dg[3] = {h->v.strintx, h->v.strinty, h->v.cb};
      for (int k = 0; k < 3; ++k)
        HIP_TRY(hipMemcpyAsync(dg[k], diag_save + (size_t)k * h->na, nA, hipMemcpyDeviceToDevice, h->stream));
    }
    return EVP1D_OK;
  }
  int take(int index) {
    for (size_t k = 0; k < used.size(); ++k)
      if (!used[k]) {
        used[k] = 1;
        RC_TRY(copy_state(slots[k], work(), false));
        return (int)k;
      }
    return -fail(EVP1D_ESTATE, "evp1d_adjoint_run: out of snapshot slots at step %d", index);
  }
  static int blocks(int n) { return (n + 255) / 256; }
  // the first reverse step: x_N from x_{N-1} (kept as the handle's final
  // state, with its str1..8, strint and Cb), the loss and the adjoint seed
  int final_step() {
    RC_TRY(copy_state(final_state, work(), true));
    DevView fv = h->v;
    fv.sig = final_state.sig;
    fv.vel = final_state.vel;
    fv.vel_out = final_state.vel;
    DevView keep = h->v;
    h->v = fv;
    int rc = launch_stress(h);
    if (!rc) rc = launch_stepu(h, true);
    if (!rc) rc = launch_halo(h);
    h->v = keep;
    RC_TRY(rc);
    const size_t nV = (size_t)h->navel * sizeof(double2);
    double2 *pairs[4] = {h->v.s15, h->v.s27, h->v.s36, h->v.s48};
    for (int k = 0; k < 4; ++k)
      HIP_TRY(hipMemcpyAsync(str_save + (size_t)k * h->navel, pairs[k], nV, hipMemcpyDeviceToDevice, h->stream));
    if (h->na > 0) {
      hipLaunchKernelGGL(k_seed, dim3(blocks(h->na)), dim3(256), 0, h->stream, fv, a, (const double *)ut,
                         (const double *)vt, terms);
      HIP_TRY(hipGetLastError());
    }
    std::vector<double> t((size_t)h->na);
    if (h->na > 0)
      HIP_TRY(hipMemcpyAsync(t.data(), terms, (size_t)h->na * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    loss = 0.0;
    for (double x : t) loss += x;
    return EVP1D_OK;
  }
  // reverse unit k with the working arrays holding its start state: one
  // subcycle from recomputed str1..8 (seg = 1), or a segment through a
  // stored velocity trajectory
  int reverse_step(int k) {
    if (seg > 1) {
      const int b = unit_start(k - 1), e = unit_start(k);
      RC_TRY(run_segment(e - b, k == nunits));
      // the segment's forward sweep left its end state (with lazy units, a
      // segment other than the last leaves the stresses stale)
      working = lazy && k != nunits ? -1 : k;
      return EVP1D_OK;
    }
    if (k == nsub) RC_TRY(final_step());
    const DevView &d = h->v;
    if (d.t_end > 0) {
      hipLaunchKernelGGL(k_str_recompute, dim3(blocks(d.t_end)), dim3(256), 0, h->stream, d);
      HIP_TRY(hipGetLastError());
    }
    if (d.u_end > d.u_begin) {
      hipLaunchKernelGGL(k_stepu_adj<false>, dim3(blocks(d.u_end - d.u_begin)), dim3(256), 0, h->stream, d, a);
      HIP_TRY(hipGetLastError());
    }
    if (d.t_end > 0) {
      hipLaunchKernelGGL(k_stress_adj<false>, dim3(blocks(d.t_end)), dim3(256), 0, h->stream, d, a);
      HIP_TRY(hipGetLastError());
    }
return EVP1D_OK;
  }
  // All nsub states kept (as much HBM as the trajectory needs): the forward
  // sweep writes each subcycle's velocities out of place into per-subcycle
  // buffers (split addressing: only the slots a subcycle changes), the
  // reverse sweep reads them back; no recomputation, no state copies.  The
  // reverse needs no stresses and no str1..8: the stress adjoint reads only
  // velocities, and the stepu adjoint takes its forward sums from u_k and
  // u_{k+1} (stepu_adjoint).
  // len subcycles from the working state, reversed: the whole run (len =
  // nsub, final) or one segment of a segmented checkpointed run.  final: the
  // segment ends at x_nsub, whose loss seeds the adjoint (and which a
  // segmented checkpointed run, the one with diag_save, saves: the sweeps of
  // earlier segments overwrite the handle's arrays).
  int run_segment(int len, bool final) {
    const DevView W = h->v;
    const int N = len, nu = W.u_end - W.u_begin, nt = W.t_end;
    // The trajectory storage stays with the handle (large allocations and
    // frees are slow) until it is destroyed or a checkpointed run releases
    // it.  Every entry is written before it is read.
    const size_t need = trajectory_bytes(h, N);
    if (h->traj_bytes < need) {
      HIP_TRY(hipStreamSynchronize(h->stream));
      if (h->traj) HIP_TRY(hipFree(h->traj));
      h->traj = nullptr;
      h->traj_bytes = 0;
      size_t free_b = 0, total_b = 0;
      HIP_TRY(hipMemGetInfo(&free_b, &total_b));
      if (need + ((size_t)1 << 30) > free_b)
        return fail(EVP1D_ENOMEM, "evp1d_adjoint_run: storing %d states needs %.1f GB, %.1f GB free; use fewer", N,
                    need / 1e9, free_b / 1e9);
      const hipError_t e = hipMalloc(&h->traj, need);
      if (e != hipSuccess) {
        h->traj = nullptr;
        return fail(EVP1D_ENOMEM, "evp1d_adjoint_run: hipMalloc(%zu): %s", need, hipGetErrorString(e));
      }
      h->traj_bytes = need;
    }
    double2 *U = static_cast<double2 *>(h->traj);
    if (nu > 0)
      HIP_TRY(hipMemcpyAsync(U, W.vel + W.u_begin, (size_t)nu * sizeof(double2), hipMemcpyDeviceToDevice, h->stream));
    // the split view of subcycle k (x_{k-1} -> x_k): velocities split, str1..8
    // in the handle's arrays (the unfused sweep's stepu reads what its stress
    // wrote there; the last subcycle leaves them as a run would)
    auto view = [&](int k) {
      DevView dk = W;
      dk.vel = U + (size_t)(k - 1) * nu;
      dk.velc = W.vel;
      dk.s15c = W.s15; dk.s27c = W.s27; dk.s36c = W.s36; dk.s48c = W.s48;
      dk.vel_out = k < N ? U + (size_t)k * nu : W.vel + W.u_begin;
      return dk;
    };
    if (pair_form(h)) {
      // two subcycles per launch, both stored (k_pair<.., TRAJ>)
      RC_TRY(ensure_fused_buffers(h));
      RC_TRY(issue_fused(h, N, U, false, sig_at, !lazy || final));
      sig_at = nullptr;
    } else if (h->fused_ok) {
      // fused forward sweep; two-buffer stresses ping-pong between the
      // handle's buffers and end in its own (subcycle k writes it when N - k
      // is even)
      RC_TRY(ensure_fused_buffers(h));
      if (!h->fused_inplace)
        HIP_TRY(hipMemcpyAsync(h->sig_alt, W.sig, (size_t)12 * h->ldsig * sizeof(double), hipMemcpyDeviceToDevice,
                               h->stream));
      FusedView f;
      RC_TRY(fused_prologue(h, W, f, h->fused_inplace != 0));
      const size_t lds = (size_t)4 * h->fused_cap * sizeof(double2);
      for (int k = 1; k <= N; ++k) {
        DevView dk = view(k);
        f.sig_out = W.sig;
        if (!h->fused_inplace) {
          const bool to_w = ((N - k) % 2) == 0;
          dk.sig = to_w ? h->sig_alt : W.sig;
          f.sig_out = to_w ? W.sig : h->sig_alt;
        }
        f.vel_out = dk.vel_out;
        fused_step(h, f, k);
        if (h->fused_ntiles > 0)
          hipLaunchKernelGGL(kFusedSplitFns[fused_block_index(h->tile_block)], dim3(h->fused_ntiles),
                             dim3(h->tile_block), lds, h->stream, dk, f, k == N ? 1 : 0);
        HIP_TRY(hipGetLastError());
      }
      RC_TRY(fused_epilogue(h, W, N, h->fused_inplace != 0));
    } else {
      for (int k = 1; k <= N; ++k) {
        const DevView dk = view(k);
        if (nt > 0) hipLaunchKernelGGL(k_stress_split, dim3(blocks(nt)), dim3(256), 0, h->stream, dk);
        if (nu > 0)
          hipLaunchKernelGGL(k_stepu_split, dim3(blocks(nu)), dim3(256), 0, h->stream, dk, k == N ? 1 : 0);
        HIP_TRY(hipGetLastError());
      }
    }
    // the handle now holds the segment's end state (stresses, velocities,
    // str1..8, strint, Cb)
    if (final) {
      if (diag_save) RC_TRY(save_final());   // a segmented checkpointed run
      if (h->na > 0) {
        hipLaunchKernelGGL(k_seed, dim3(blocks(h->na)), dim3(256), 0, h->stream, W, a, (const double *)ut,
                           (const double *)vt, terms);
        HIP_TRY(hipGetLastError());
      }
      std::vector<double> t((size_t)h->na);
      if (h->na > 0)
        HIP_TRY(hipMemcpyAsync(t.data(), terms, (size_t)h->na * sizeof(double), hipMemcpyDeviceToHost, h->stream));
      HIP_TRY(hipStreamSynchronize(h->stream));
      loss = 0.0;
      for (double x : t) loss += x;
    }
    // the reverse kernels take the handle's cache hints (pair_hints)
    const bool hints = pair_hints(h);
    const auto stepu_adj = hints ? k_stepu_adj<true, true> : k_stepu_adj<true, false>;
    const auto stress_adj = hints ? k_stress_adj<true, true> : k_stress_adj<true, false>;
    for (int k = N; k >= 1; --k) {
      const DevView dk = view(k);
      if (nu > 0) hipLaunchKernelGGL(stepu_adj, dim3(blocks(nu)), dim3(256), 0, h->stream, dk, a);
      if (nt > 0) hipLaunchKernelGGL(stress_adj, dim3(blocks(nt)), dim3(256), 0, h->stream, dk, a);
      HIP_TRY(hipGetLastError());
    }
    return EVP1D_OK;
  }
  // reverse steps a+1 .. a+n; x_a is stored in slot sa; s further free slots
  int rev(int a0, int n, int s, int sa) {
    if (n == 1) {
      RC_TRY(restore(sa, a0));
      return reverse_step(a0 + 1);
    }
    if (s == 0) {
      for (int i = n - 1; i >= 0; --i) {
        RC_TRY(restore(sa, a0));
        RC_TRY(advance(i));
        RC_TRY(reverse_step(a0 + i + 1));
      }
      return EVP1D_OK;
    }
    const int m = sched.M(n, s);
    RC_TRY(restore(sa, a0));
    RC_TRY(advance(m));
    const int sb = take(a0 + m);
    if (sb < 0) return -sb;
    RC_TRY(rev(a0 + m, n - m, s - 1, sb));
    used[(size_t)sb] = 0;
    return rev(a0, m, s, sa);
  }
};

static int upload_caller_order(evp1d_handle *h, double *dst, const double *src) {
  if (h->na == 0) return EVP1D_OK;
  HIP_TRY(hipMemcpyAsync(h->scratch, src, (size_t)h->na * sizeof(double), hipMemcpyHostToDevice, h->stream));
  const Span sp{0, h->na, h->na, h->na};
  hipLaunchKernelGGL(k_scatter_plain, dim3(blocks_for(h->na)), dim3(kBlock), 0, h->stream, dst,
                     (const double *)h->scratch, (const int32_t *)h->inv_dev, h->na, sp);
  HIP_TRY(hipGetLastError());
  return EVP1D_OK;
}

static int download_caller_order(evp1d_handle *h, double *dst, const double *src) {
  if (!dst || h->na == 0) return EVP1D_OK;
  const Span sp{0, h->na, h->na, h->na};
  hipLaunchKernelGGL(k_gather_plain, dim3(blocks_for(h->na)), dim3(kBlock), 0, h->stream, h->scratch, src,
                     (const int32_t *)h->inv_dev, h->na, sp);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(dst, h->scratch, (size_t)h->na * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

int evp1d_adjoint_run(evp1d_handle *h, int32_t nsub, int32_t nsnap, const double *u_target, const double *v_target,
                      double *loss, double *g_strength, double *g_cdn_ocn, double *g_forcex, double *g_forcey) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_adjoint_run: NULL handle");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_adjoint_run: neighbour lists not set");
  if (nsub < 1) return fail(EVP1D_EINVAL, "evp1d_adjoint_run: nsub=%d < 1", nsub);
  if (h->na > 0 && (!u_target || !v_target)) return fail(EVP1D_EINVAL, "evp1d_adjoint_run: NULL target");
  if (h->v.nhalo > 0)
    return fail(EVP1D_ESTATE, "evp1d_adjoint_run: a ghost copy list is set; its reverse is not implemented");
  RC_TRY(ensure(h));
  AdjointRun R;
  R.h = h;
  R.nsub = nsub;
  const size_t nA = (size_t)h->na * sizeof(double), nV = (size_t)h->navel * sizeof(double2);
  int *cnz = nullptr;
  // adjoint state, gradient accumulators, targets and loss terms: carved
  // from one allocation that the handle keeps (large allocations are slow),
  // cleared on every call
  {
    const size_t parts[12] = {nV, R.sig_bytes(), nV, 4 * (size_t)h->na * sizeof(double2), nA, nA, nA, nA,
                              nA, nA, nA, sizeof(int)};
    size_t total = 0;
    for (size_t b : parts) total += (b + 255) / 256 * 256;
    if (h->adj_ws_bytes < total) {
      HIP_TRY(hipStreamSynchronize(h->stream));
      if (h->adj_ws) HIP_TRY(hipFree(h->adj_ws));
      h->adj_ws = nullptr;
      h->adj_ws_bytes = 0;
      const hipError_t e = hipMalloc(&h->adj_ws, total);
      if (e != hipSuccess) {
        h->adj_ws = nullptr;
        return fail(EVP1D_ENOMEM, "evp1d_adjoint_run: hipMalloc(%zu): %s", total, hipGetErrorString(e));
      }
      h->adj_ws_bytes = total;
    }
    HIP_TRY(hipMemsetAsync(h->adj_ws, 0, total, h->stream));
    char *base = static_cast<char *>(h->adj_ws);
    size_t off = 0;
    int part = 0;
    auto carve = [&]() {
      void *p = base + off;
      off += (parts[part++] + 255) / 256 * 256;
      return p;
    };
    R.a.ub = static_cast<double2 *>(carve());
    R.a.sigb = static_cast<double *>(carve());
    R.a.g = static_cast<double2 *>(carve());
    R.a.vc = static_cast<double2 *>(carve());
    R.a.gstr = static_cast<double *>(carve());
    R.a.gcw = static_cast<double *>(carve());
    R.a.gfx = static_cast<double *>(carve());
    R.a.gfy = static_cast<double *>(carve());
    R.ut = static_cast<double *>(carve());
    R.vt = static_cast<double *>(carve());
    R.terms = static_cast<double *>(carve());
    cnz = static_cast<int *>(carve());
  }
  R.a.cnz = cnz;
  if (h->navel > h->v.t_end) {
    hipLaunchKernelGGL(k_const_flag, dim3((unsigned)((h->navel - h->v.t_end + 255) / 256)), dim3(256), 0, h->stream,
                       h->v, cnz);
    HIP_TRY(hipGetLastError());
  }
  // constant fields and stress neighbours in 16-byte groups for the
  // reverse kernels (the same buffers the pair form packs per run)
  if (!h->pconst) HIP_TRY(hipMalloc((void **)&h->pconst, (size_t)kPairConst * std::max(h->na, 1) * sizeof(double2)));
  if (!h->nbr4) HIP_TRY(hipMalloc((void **)&h->nbr4, (size_t)std::max(h->na, 1) * sizeof(int4)));
  if (h->na > 0) {
    hipLaunchKernelGGL(k_pack_const, dim3(blocks_for(h->na)), dim3(kBlock), 0, h->stream, h->v, h->pconst,
                       (long)h->na);
    HIP_TRY(hipGetLastError());
  }
  if (h->v.t_end > 0) {
    hipLaunchKernelGGL(k_pack_nbr, dim3(blocks_for(h->v.t_end)), dim3(kBlock), 0, h->stream, h->v, h->nbr4,
                       (long)h->v.t_end);
    HIP_TRY(hipGetLastError());
  }
  h->pconst_fresh = true;   // the forward sweeps of this call reuse the packed fields
  R.a.pc = h->pconst;
  R.a.lp = h->na;
  R.a.nbr = h->nbr4;
  RC_TRY(upload_caller_order(h, R.ut, u_target));
  RC_TRY(upload_caller_order(h, R.vt, v_target));
  // gradient storage the handle holds from earlier calls: reused, or given
  // back when a call needs the room
  auto release = [&](void *&p, size_t &bytes) -> int {
    if (!p) return EVP1D_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipFree(p));
    p = nullptr;
    bytes = 0;
    return EVP1D_OK;
  };
  const size_t held = h->traj_bytes + h->ckpt_bytes;
  if (nsnap <= 0) {
    // automatic: the whole trajectory if it fits in free device memory, else
    // binomial checkpointing with as many stored states as fit
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const size_t reserve = (size_t)1 << 30;
    const size_t traj = trajectory_bytes(h, nsub);
    const size_t have = h->traj_bytes >= traj ? traj : 0;   // storage the handle already holds
    if (have > 0 || traj + reserve <= free_b + held) {
      nsnap = nsub;
    } else {
      const size_t slot = R.sig_bytes() + (size_t)h->na * sizeof(double2);
      const size_t room = free_b + held > reserve ? (free_b + held - reserve) / slot : 0;
      nsnap = (int)std::max<size_t>(1, std::min<size_t>(room / 2, (size_t)nsub - 1));
    }
  }
  if (nsnap >= nsub) {
    // a whole-trajectory run that has to allocate its storage gives the
    // stored states of earlier checkpointed runs back first
    if (h->traj_bytes < trajectory_bytes(h, nsub)) RC_TRY(release(h->ckpt, h->ckpt_bytes));
    R.seg = nsub;
    R.nunits = 1;
    RC_TRY(R.run_segment(nsub, true));
  } else {
    // Checkpointing units of seg subcycles: binomial checkpointing over the
    // units, each unit reversed through a stored velocity trajectory of seg
    // states (seg = 1: one subcycle from recomputed str1..8).  Automatic:
    // the shortest unit that gives every unit start a stored state (no
    // recomputation beyond one forward sweep per unit), if its trajectory
    // fits beside the stored states, else seg = 1.
    const int S0 = nsnap;
    int seg = h->adj_segment > 0 ? std::min<int>(h->adj_segment, nsub) : (nsub + S0 - 1) / S0;
    if (h->adj_segment == 0 && seg > 1) {
      size_t free_b = 0, total_b = 0;
      HIP_TRY(hipMemGetInfo(&free_b, &total_b));
      const size_t slot = R.sig_bytes() + (size_t)h->na * sizeof(double2);
      if (trajectory_bytes(h, seg) + (size_t)S0 * slot + ((size_t)1 << 30) > free_b + held) seg = 1;
    }
    // a checkpointed run keeps no more trajectory storage than its units need
    // (a whole trajectory of an earlier call is given back)
    if (h->traj_bytes > (seg > 1 ? trajectory_bytes(h, seg) : 0)) RC_TRY(release(h->traj, h->traj_bytes));
    R.seg = seg;
    R.nunits = (nsub + seg - 1) / seg;
    R.lazy = pair_form(h) && h->fused_ok && seg > 1 && seg % 2 == 0 && (nsub - (R.nunits - 1) * seg) % 2 == 0;
    const int S = std::min(S0, R.nunits);   // more slots than units are never used
    RC_TRY(R.carve_checkpoints(S));
    R.sched.build(R.nunits, S - 1);
    R.working = 0;
    const int s0 = R.take(0);
    if (s0 < 0) return -s0;
    RC_TRY(R.rev(0, R.nunits, S - 1, s0));
    // leave the handle in the forward final state x_N
    RC_TRY(R.restore_final());
  }
  RC_TRY(download_caller_order(h, g_strength, R.a.gstr));
  RC_TRY(download_caller_order(h, g_cdn_ocn, R.a.gcw));
  RC_TRY(download_caller_order(h, g_forcex, R.a.gfx));
  RC_TRY(download_caller_order(h, g_forcey, R.a.gfy));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (loss) *loss = R.loss;
  return EVP1D_OK;
}

int evp1d_tangent_run(evp1d_handle *h, int32_t nsub, const double *d_strength, const double *d_cdn_ocn,
                      const double *d_forcex, const double *d_forcey, const double *u_target,
                      const double *v_target, double *loss, double *dloss) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_tangent_run: NULL handle");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_tangent_run: neighbour lists not set");
  if (nsub < 0) return fail(EVP1D_EINVAL, "evp1d_tangent_run: nsub=%d < 0", nsub);
  if (h->na > 0 && (!u_target || !v_target)) return fail(EVP1D_EINVAL, "evp1d_tangent_run: NULL target");
  if (h->v.nhalo > 0) return fail(EVP1D_ESTATE, "evp1d_tangent_run: a ghost copy list is set");
  RC_TRY(ensure(h));
  const size_t nA = (size_t)h->na * sizeof(double), nV = (size_t)h->navel * sizeof(double2);
  std::vector<void *> mem;
  struct Free {
    std::vector<void *> &m;
    hipStream_t s;
    ~Free() {
      (void)hipStreamSynchronize(s);
      for (void *p : m) (void)hipFree(p);
    }
  } guard{mem, h->stream};
  auto A = [&](void **p, size_t bytes) -> int {
    const hipError_t e = hipMalloc(p, bytes ? bytes : 256);
    if (e != hipSuccess) return fail(EVP1D_ENOMEM, "evp1d_tangent_run: hipMalloc(%zu)", bytes);
    mem.push_back(*p);
    HIP_TRY(hipMemsetAsync(*p, 0, bytes ? bytes : 256, h->stream));
    return EVP1D_OK;
  };
  TanView t{};
  double *dir[4] = {nullptr, nullptr, nullptr, nullptr}, *ut = nullptr, *vt = nullptr, *terms = nullptr;
  RC_TRY(A((void **)&t.dsig, (size_t)12 * h->ldsig * sizeof(double)));
  RC_TRY(A((void **)&t.dvel, nV));
  RC_TRY(A((void **)&t.ds15, nV));
  RC_TRY(A((void **)&t.ds27, nV));
  RC_TRY(A((void **)&t.ds36, nV));
  RC_TRY(A((void **)&t.ds48, nV));
  const double *src[4] = {d_strength, d_cdn_ocn, d_forcex, d_forcey};
  for (int k = 0; k < 4; ++k) {
    RC_TRY(A((void **)&dir[k], nA));
    if (src[k]) RC_TRY(upload_caller_order(h, dir[k], src[k]));
  }
  t.dstrength = dir[0];
  t.dcw = dir[1];
  t.dfx = dir[2];
  t.dfy = dir[3];
  RC_TRY(A((void **)&ut, nA));
  RC_TRY(A((void **)&vt, nA));
  RC_TRY(A((void **)&terms, 2 * nA));
  RC_TRY(upload_caller_order(h, ut, u_target));
  RC_TRY(upload_caller_order(h, vt, v_target));
  const DevView &d = h->v;
  for (int k = 0; k < nsub; ++k) {
    if (d.t_end > 0) hipLaunchKernelGGL(k_stress_tl, dim3((d.t_end + 255) / 256), dim3(256), 0, h->stream, d, t);
    if (d.u_end > d.u_begin)
      hipLaunchKernelGGL(k_stepu_tl, dim3((d.u_end - d.u_begin + 255) / 256), dim3(256), 0, h->stream, d, t,
                         k == nsub - 1 ? 1 : 0);
    HIP_TRY(hipGetLastError());
  }
  double l = 0.0, dl = 0.0;
  if (h->na > 0) {
    hipLaunchKernelGGL(k_loss_tl, dim3(blocks_for(h->na)), dim3(kBlock), 0, h->stream, d, t, (const double *)ut,
                       (const double *)vt, terms);
    HIP_TRY(hipGetLastError());
    std::vector<double> buf(2 * (size_t)h->na);
    HIP_TRY(hipMemcpyAsync(buf.data(), terms, 2 * nA, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int q = 0; q < h->na; ++q) {
      l += buf[(size_t)q];
      dl += buf[(size_t)h->na + q];
    }
  }
  if (loss) *loss = l;
  if (dloss) *dloss = dl;
  return EVP1D_OK;
}

int evp1d_loss(evp1d_handle *h, const double *u_target, const double *v_target, double *loss) {
  if (!h || !loss) return fail(EVP1D_EINVAL, "evp1d_loss: NULL argument");
  if (h->na > 0 && (!u_target || !v_target)) return fail(EVP1D_EINVAL, "evp1d_loss: NULL target");
  RC_TRY(ensure(h));
  *loss = 0.0;
  if (h->na == 0) return EVP1D_OK;
  double *buf = nullptr;
  HIP_TRY(hipMalloc(&buf, 3 * (size_t)h->na * sizeof(double)));
  int rc = upload_caller_order(h, buf, u_target);
  if (!rc) rc = upload_caller_order(h, buf + h->na, v_target);
  if (!rc) {
    AdjView none{};
    hipLaunchKernelGGL(k_seed, dim3(blocks_for(h->na)), dim3(kBlock), 0, h->stream, h->v, none,
                       (const double *)buf, (const double *)(buf + h->na), buf + 2 * h->na);
    if (hipGetLastError() != hipSuccess) rc = fail(EVP1D_EHIP, "evp1d_loss: launch failed");
  }
  std::vector<double> t((size_t)h->na);
  if (!rc && hipMemcpyAsync(t.data(), buf + 2 * h->na, (size_t)h->na * sizeof(double), hipMemcpyDeviceToHost,
                            h->stream) != hipSuccess)
    rc = fail(EVP1D_EHIP, "evp1d_loss: copy failed");
  if (hipStreamSynchronize(h->stream) != hipSuccess && !rc) rc = fail(EVP1D_EHIP, "evp1d_loss: sync failed");
  (void)hipFree(buf);
  if (rc) return rc;
  for (double x : t) *loss += x;
  return EVP1D_OK;
}

// ---------------------------------------------------------------------------
// evp1d_loss_noalloc: evp1d_loss with a fixed device scratch array instead of
// per-call allocations.  `scratch` holds the two targets in caller order
// (kLossMaxPoints doubles each), then one partial sum per block of the
// reduction, then the result.  Host code never passes `scratch` itself to a
// kernel or a copy: each device's own address of it comes from
// hipGetSymbolAddress (scratch_ptr) and only that address is used.  Every call
// on a device shares that device's array; calls that overlap on one device
// race.
// ---------------------------------------------------------------------------
constexpr long kLossMaxPoints = 1L << 22;   // 4,194,304 points: 64 MiB of targets
constexpr int kLossBlock = 256;
constexpr long kLossMaxBlocks = (kLossMaxPoints + kLossBlock - 1) / kLossBlock;
constexpr long kLossScratch = 2 * kLossMaxPoints + kLossMaxBlocks + 1;
static __device__ __align__(256) double scratch[kLossScratch];

constexpr int kMaxDevices = 64;
static double *g_scratch_dev[kMaxDevices];   // per device: the address of its `scratch`

// Look up the current device's address of `scratch` once (the caller has
// made device `dev` current).  quiet: leave the error message alone.
static int scratch_setup(int dev, bool quiet) {
  if (dev < 0 || dev >= kMaxDevices)
    return quiet ? EVP1D_EINVAL : fail(EVP1D_EINVAL, "scratch: device %d out of range", dev);
  if (__atomic_load_n(&g_scratch_dev[dev], __ATOMIC_ACQUIRE)) return EVP1D_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return quiet ? EVP1D_EHIP
                 : fail(EVP1D_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
  __atomic_store_n(&g_scratch_dev[dev], static_cast<double *>(p), __ATOMIC_RELEASE);
  return EVP1D_OK;
}

// per block: the sum of the loss terms of its slots, in a fixed tree order;
// the target of slot s is the caller's element perm[s]
__global__ __launch_bounds__(kLossBlock) void k_loss_partial(DevView d, const int32_t *perm, const double *ut,
                                                              const double *vt, double *partial) {
  __shared__ double red[kLossBlock];
  const int s = blockIdx.x * kLossBlock + threadIdx.x;
  double t = 0.0;
  if (s >= d.u_begin && s < d.u_end && s < d.na) {
    const double2 v = d.vel[s];
    const int q = perm[s];
    const double ex = v.x - ut[q], ey = v.y - vt[q];
    t = 0.5 * (ex * ex + ey * ey);
  }
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = kLossBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// one block: out[0] = the sum of partial[0..n) (each thread a strided run,
// then a tree)
__global__ __launch_bounds__(kLossBlock) void k_loss_final(const double *partial, long n, double *out) {
  __shared__ double red[kLossBlock];
  double t = 0.0;
  for (long k = threadIdx.x; k < n; k += kLossBlock) t += partial[k];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = kLossBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

int evp1d_loss_noalloc(evp1d_handle *h, const double *u_target, const double *v_target, double *loss) {
  if (!h || !loss) return fail(EVP1D_EINVAL, "evp1d_loss_noalloc: NULL argument");
  if (h->na > kLossMaxPoints)
    return fail(EVP1D_EINVAL, "evp1d_loss_noalloc: na = %d exceeds the fixed scratch of %ld points; use evp1d_loss",
                h->na, kLossMaxPoints);
  if (h->na > 0 && (!u_target || !v_target)) return fail(EVP1D_EINVAL, "evp1d_loss_noalloc: NULL target");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_loss_noalloc: neighbour lists not set");
  // bringing the layout up to date may allocate: that is left to the other calls
  if (h->layout_dirty || h->idx_dirty || h->perm_dev_dirty)
    return fail(EVP1D_ESTATE, "evp1d_loss_noalloc: the slot layout is out of date; run evp1d_run or evp1d_loss "
                              "once after changing lists or skip flags");
  *loss = 0.0;
  if (h->na == 0) return EVP1D_OK;
  RC_TRY(scratch_setup(h->device, false));
  double *sp = __atomic_load_n(&g_scratch_dev[h->device], __ATOMIC_ACQUIRE);
  double *ut = sp, *vt = sp + kLossMaxPoints, *partial = sp + 2 * kLossMaxPoints;
  double *out = partial + kLossMaxBlocks;
  const size_t bytes = (size_t)h->na * sizeof(double);
  HIP_TRY(hipMemcpyAsync(ut, u_target, bytes, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(vt, v_target, bytes, hipMemcpyHostToDevice, h->stream));
  const long nb = (h->na + kLossBlock - 1) / kLossBlock;
  hipLaunchKernelGGL(k_loss_partial, dim3((unsigned)nb), dim3(kLossBlock), 0, h->stream, h->v,
                     (const int32_t *)h->perm_dev, (const double *)ut, (const double *)vt, partial);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(kLossBlock), 0, h->stream, (const double *)partial, nb, out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(loss, out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

static int ensure_pos2d(evp1d_handle *h) {
  if (!h->pos2d_dirty && h->pos2d_dev) return EVP1D_OK;
  if (!h->pos2d_dev) HIP_TRY(hipMalloc((void **)&h->pos2d_dev, (size_t)std::max(h->na, 1) * sizeof(int32_t)));
  std::vector<int32_t> pos((size_t)h->na);
  for (int s = 0; s < h->na; ++s) pos[(size_t)s] = h->grid_ij[(size_t)h->perm[(size_t)s]];
  if (h->na > 0)
    HIP_TRY(hipMemcpyAsync(h->pos2d_dev, pos.data(), (size_t)h->na * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->pos2d_dirty = false;
  return EVP1D_OK;
}

int evp1d_set_grid(evp1d_handle *h, int32_t nx, int32_t ny, const int32_t *indxtij, int32_t n) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_grid: NULL handle");
  if (h->m_na.size() != 1) return fail(EVP1D_EINVAL, "evp1d_set_grid: only for single-domain handles");
  if (nx < 1 || ny < 1 || (long)nx * ny > 0x7fffffffL) return fail(EVP1D_EINVAL, "evp1d_set_grid: bad nx=%d ny=%d", nx, ny);
  if (n != h->navel) return fail(EVP1D_EINVAL, "evp1d_set_grid: length %d != navel %d", n, h->navel);
  if (n > 0 && !indxtij) return fail(EVP1D_EINVAL, "evp1d_set_grid: indxtij is NULL");
  std::vector<int32_t> ij((size_t)h->na);
  for (int q = 0; q < h->na; ++q) {
    const int32_t v = indxtij[q];
    if (v < 1 || v > nx * ny)
      return fail(EVP1D_EINDEX, "evp1d_set_grid: indxtij(%d) = %d outside 1..%d", q + 1, v, nx * ny);
    ij[(size_t)q] = v - 1;
  }
  if (nx != h->grid_nx || ny != h->grid_ny)
    for (double *&g : h->g2d) {
      if (g) HIP_TRY(hipFree(g));
      g = nullptr;
    }
  h->grid_nx = nx;
  h->grid_ny = ny;
  h->grid_ij.swap(ij);
  h->pos2d_dirty = true;
  return EVP1D_OK;
}

// wait = false: evp1d_put_2d_async (the copies stay queued)
static int put_2d_impl(evp1d_handle *h, int32_t nfield, const int32_t *fields, const double *arrays, bool wait,
                       const char *name) {
  if (!h) return fail(EVP1D_EINVAL, "%s: NULL handle", name);
  if (h->grid_nx == 0) return fail(EVP1D_ESTATE, "%s: evp1d_set_grid was not called", name);
  if (nfield < 0 || (nfield > 0 && (!fields || !arrays))) return fail(EVP1D_EINVAL, "%s: bad arguments", name);
  const size_t n2 = (size_t)h->grid_nx * h->grid_ny;
  FieldRef r;
  for (int k = 0; k < nfield; ++k)
    if (!field_ref(h, fields[k], r)) return fail(EVP1D_EINVAL, "%s: unknown field %d", name, fields[k]);
  RC_TRY(ensure_pos2d(h));
  const int na = h->na, navel = h->navel;
  for (int k = 0; k < nfield; ++k) {
    const int f = fields[k];
    field_ref(h, f, r);
    if (!h->g2d[f]) HIP_TRY(hipMalloc((void **)&h->g2d[f], n2 * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(h->g2d[f], arrays + (size_t)k * n2, n2 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (r.kind == FK_PLAIN) {
      if (na > 0)
        hipLaunchKernelGGL(k_pack2d_plain, dim3(blocks_for(na)), dim3(kBlock), 0, h->stream, r.plain,
                           (const double *)h->g2d[f], (const int32_t *)h->pos2d_dev, na);
    } else {
      const int zg = (f == EVP1D_UVEL || f == EVP1D_VVEL) ? 1 : 0;   // ice_dyn_evp1d.F90:880-881
      if (navel > 0)
        hipLaunchKernelGGL(k_pack2d_comp, dim3(blocks_for(navel)), dim3(kBlock), 0, h->stream, r.pair,
                           (const double *)h->g2d[f], (const int32_t *)h->pos2d_dev, na, navel, r.comp, zg);
    }
    HIP_TRY(hipGetLastError());
  }
  if (wait) HIP_TRY(hipStreamSynchronize(h->stream));   // the caller may reuse its arrays
  return EVP1D_OK;
}

int evp1d_put_2d(evp1d_handle *h, int32_t nfield, const int32_t *fields, const double *arrays) {
  return put_2d_impl(h, nfield, fields, arrays, true, "evp1d_put_2d");
}

int evp1d_put_2d_async(evp1d_handle *h, int32_t nfield, const int32_t *fields, const double *arrays) {
  return put_2d_impl(h, nfield, fields, arrays, false, "evp1d_put_2d_async");
}

int evp1d_get_2d(evp1d_handle *h, int32_t nfield, const int32_t *fields, double *arrays) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_2d: NULL handle");
  if (h->grid_nx == 0) return fail(EVP1D_ESTATE, "evp1d_get_2d: evp1d_set_grid was not called");
  if (nfield < 0 || (nfield > 0 && (!fields || !arrays))) return fail(EVP1D_EINVAL, "evp1d_get_2d: bad arguments");
  const size_t n2 = (size_t)h->grid_nx * h->grid_ny;
  FieldRef r;
  for (int k = 0; k < nfield; ++k) {
    if (!field_ref(h, fields[k], r)) return fail(EVP1D_EINVAL, "evp1d_get_2d: unknown field %d", fields[k]);
    if (!h->g2d[fields[k]])
      return fail(EVP1D_ESTATE, "evp1d_get_2d: field %d was never put with evp1d_put_2d", fields[k]);
  }
  RC_TRY(ensure_pos2d(h));
  const int na = h->na;
  for (int k = 0; k < nfield; ++k) {
    const int f = fields[k];
    field_ref(h, f, r);
    if (na > 0) {
      if (r.kind == FK_PLAIN)
        hipLaunchKernelGGL(k_unpack2d_plain, dim3(blocks_for(na)), dim3(kBlock), 0, h->stream, h->g2d[f],
                           (const double *)r.plain, (const int32_t *)h->pos2d_dev, na);
      else
        hipLaunchKernelGGL(k_unpack2d_comp, dim3(blocks_for(na)), dim3(kBlock), 0, h->stream, h->g2d[f],
                           (const double2 *)r.pair, (const int32_t *)h->pos2d_dev, na, r.comp);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(arrays + (size_t)k * n2, h->g2d[f], n2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

int evp1d_host_alloc(int64_t bytes, void **ptr) {
  if (!ptr || bytes < 0) return fail(EVP1D_EINVAL, "evp1d_host_alloc: bad arguments");
  *ptr = nullptr;
  const hipError_t e = hipHostMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1), hipHostMallocDefault);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(EVP1D_ENOMEM, "evp1d_host_alloc(%lld): %s", (long long)bytes, hipGetErrorString(e));
  }
  return EVP1D_OK;
}

int evp1d_host_free(void *ptr) {
  if (ptr) HIP_TRY(hipHostFree(ptr));
  return EVP1D_OK;
}

int evp1d_synchronize(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_synchronize: NULL handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

int evp1d_timer_start(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_timer_start: NULL handle");
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  return EVP1D_OK;
}

int evp1d_timer_stop(evp1d_handle *h, float *ms) {
  if (!h || !ms) return fail(EVP1D_EINVAL, "evp1d_timer_stop: NULL argument");
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  HIP_TRY(hipEventElapsedTime(ms, h->ev0, h->ev1));
  return EVP1D_OK;
}

int evp1d_last_error(char *buf, int32_t n) {
  if (buf && n > 0) {
    std::strncpy(buf, g_err_msg, (size_t)n - 1);
    buf[n - 1] = '\0';
  }
  return g_err_code;
}

int evp1d_build_info(int32_t *abi_version, int32_t *fma) {
  if (abi_version) *abi_version = EVP1D_ABI_VERSION;
  if (fma) *fma = EVP1D_FMA;
  return EVP1D_OK;
}

}  // extern "C"
