// evp2d_device.hip - the v0 2D block path: gfx950 kernels and its C ABI.
//
// The reference's v0 benchmark (src/evp_v0.F90, src/bench_v0.F90) keeps
// CICE's own layout: every field is an (nx_block, ny_block, max_blocks)
// array, and per block the kernels walk a list of T cells (icellt, indxti,
// indxtj) for stress and of U cells (icellu, indxui, indxuj) for stepu; with
// more than one block a copy-list halo update of uvel and vvel follows every
// subcycle (src/vars.F90:100-143).  Here the same arrays live on the device:
//
//   vel, vinit  double2[n]   (uvel, vvel) and (uvel_init, vvel_init)
//   sig         double[12][n] stressp_1..4, stressm_1..4, stress12_1..4
//   s15..s48    double2[n]   str(:,:,1/5), (2/7), (3/6), (4/8): the pairs stepu
//                            reads at (i,j), (i+1,j), (i,j+1), (i+1,j+1)
//   the rest    double[n]    one array per field
//
// n = max_blocks * nx_block * ny_block, cell (i, j, b) at linear index
// (b-1)*nx_block*ny_block + (j-1)*nx_block + (i-1), Fortran order.  The cell
// lists of all blocks are concatenated into one list of linear indices, so
// one launch covers every block: the neighbours of a cell are at fixed
// offsets (-1, -nx_block, ...), so every access is a stencil access and
// consecutive list entries of a row touch consecutive addresses.
//
// The reference clears str (all 8 planes of the block) at the start of every
// stress call (bench_v0.F90:135) and then writes it at the list's T cells.
// Here str is cleared once when the lists change and written at the T cells
// only: the cells outside the list are never written, so they stay zero and
// every read gives what the reference reads.  (The reference keeps one block
// of str scratch for all blocks; here each block has its own, which is what
// every block's stepu reads in the reference's serial loop.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/evp1d.h"
#include "evp1d_cache.hpp"
#include "evp1d_point.hpp"
#include "evp1d_tiles.hpp"
#include "evp1d_xcd.hpp"

extern "C" int evp1d__set_error(int code, const char *msg);

namespace {

int fail2(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return evp1d__set_error(code, buf);
}

#define HIP2(expr)                                                                      \
  do {                                                                                  \
    const hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return fail2(EVP1D_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
#define RC2(expr)           \
  do {                      \
    const int r_ = (expr);  \
    if (r_) return r_;      \
  } while (0)

evp1d::Scalars scalars2(const evp1d_params &p) {
  evp1d::Scalars s;
  s.capping = p.capping; s.e_factor = p.e_factor; s.epp2i = p.epp2i;
  s.deltaminEVP = p.deltaminEVP; s.arlx1i = p.arlx1i; s.denom1 = p.denom1;
  s.brlx = p.brlx; s.revp = p.revp; s.Ktens = p.Ktens; s.u0 = p.u0;
  s.cosw = p.cosw; s.sinw = p.sinw; s.rhow = p.rhow;
  return s;
}

// plain double fields, in the order of their EVP2D_* ids that are plain
enum Plain {
  P_STRENGTH, P_DXT, P_DYT, P_DXHY, P_DYHX, P_CXP, P_CYP, P_CXM, P_CYM, P_DMIN, P_UAREAR, P_CW, P_AIX,
  P_UOCN, P_VOCN, P_WATERX, P_WATERY, P_FORCEX, P_FORCEY, P_UMASSDTI, P_FM, P_STRINTX, P_STRINTY, P_TBU,
  P_TAUBX, P_TAUBY, P_COUNT
};

struct View2 {
  long n;
  int nxb;
  double2 *vel;
  const double2 *vinit;
  double *sig;
  double2 *s15, *s27, *s36, *s48;
  const double *dxt, *dyt, *dxhy, *dyhx, *cxp, *cyp, *cxm, *cym, *dmin, *strength;
  const double *cw, *aix, *uocn, *vocn, *waterx, *watery, *forcex, *forcey, *umassdti, *fm, *uarear, *tbu;
  double *strintx, *strinty, *taubx, *tauby;
  const int32_t *tcell, *ucell;
  int nt, nu;
  evp1d::Scalars s;
};

constexpr int kB2 = 256;

// stress at the T cells of every block (bench_v0.F90:137-329)
__global__ __launch_bounds__(kB2) void k2_stress(View2 v) {
  const int k = blockIdx.x * kB2 + threadIdx.x;
  if (k >= v.nt) return;
  const long c = v.tcell[k];
  const double2 cc = v.vel[c];               // (i, j)
  const double2 ve = v.vel[c - 1];           // (i-1, j)
  const double2 vs = v.vel[c - v.nxb];       // (i, j-1)
  const double2 vn = v.vel[c - v.nxb - 1];   // (i-1, j-1)
  double sig[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) sig[q] = v.sig[q * v.n + c];
  evp1d::StressOut o;
  evp1d::stress_core(v.s, cc.x, cc.y, ve.x, ve.y, vs.x, vs.y, vn.x, vn.y, v.dxt[c], v.dyt[c], v.strength[c],
                     v.cxp[c], v.cyp[c], v.cxm[c], v.cym[c], v.dmin[c], v.dxhy[c], v.dyhx[c], sig, o);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v.sig[q * v.n + c] = o.sp[q];
    v.sig[(4 + q) * v.n + c] = o.sm[q];
    v.sig[(8 + q) * v.n + c] = o.s12[q];
  }
  v.s15[c] = make_double2(o.str[0], o.str[4]);
  v.s27[c] = make_double2(o.str[1], o.str[6]);
  v.s36[c] = make_double2(o.str[2], o.str[5]);
  v.s48[c] = make_double2(o.str[3], o.str[7]);
}

// stepu at the U cells of every block (bench_v0.F90:563-600)
__global__ __launch_bounds__(kB2) void k2_stepu(View2 v) {
  const int k = blockIdx.x * kB2 + threadIdx.x;
  if (k >= v.nu) return;
  const long c = v.ucell[k];
  const double2 vv = v.vel[c];
  const double2 vi = v.vinit[c];
  const double2 a = v.s15[c];                // str(i,j,1), str(i,j,5)
  const double2 b = v.s27[c + 1];            // str(i+1,j,2), str(i+1,j,7)
  const double2 e = v.s36[c + v.nxb];        // str(i,j+1,3), str(i,j+1,6)
  const double2 f = v.s48[c + v.nxb + 1];    // str(i+1,j+1,4), str(i+1,j+1,8)
  evp1d::StepuOut o;
  evp1d::stepu_point(v.s, vv.x, vv.y, v.cw[c], v.aix[c], v.uocn[c], v.vocn[c], v.waterx[c], v.watery[c],
                     v.forcex[c], v.forcey[c], v.umassdti[c], v.fm[c], v.uarear[c], vi.x, vi.y, v.tbu[c],
                     a.x, b.x, e.x, f.x, a.y, e.y, b.y, f.y, o);
  v.vel[c] = make_double2(o.uvel, o.vvel);
  v.strintx[c] = o.strintx;
  v.strinty[c] = o.strinty;
  v.taubx[c] = -(o.uvel * o.Cb);
  v.tauby[c] = -(o.vvel * o.Cb);
}

// Fused subcycle: stress and stepu of one tile of kTW x kTH cells per
// workgroup.  The tile's T cells, plus the T cells of the column east of it
// and the row north of it that its U cells read (evaluated again, not
// stored), put their str1..8 into LDS tables laid out like the cells,
// (kTW+1) x (kTH+1) entries, zero where no T cell is (the reference's
// cleared str); after a barrier stepu reads its four entries at (i,j),
// (i+1,j), (i,j+1), (i+1,j+1) by position.  Velocities and stresses are read
// from the view and written to f.vel_out / f.sig_out, the other buffer, so no
// tile reads a half-updated subcycle.  str1..8, strintx/y and taubx/y are
// observable after the run only: stored in the last subcycle (`last`).
#ifndef EVP2D_TILE_W
#define EVP2D_TILE_W 32
#endif
#ifndef EVP2D_TILE_H
#define EVP2D_TILE_H 8
#endif
constexpr int kTW = EVP2D_TILE_W, kTH = EVP2D_TILE_H, kCap2 = (kTW + 1) * (kTH + 1);
constexpr int kB2F = 256;
struct Tile2 {
  int32_t t0, nto, nth;   // T entries [t0, t0 + nto) owned, then nth halo entries
  int32_t u0, nu;         // U entries [u0, u0 + nu)
};
// During a fused run the per-cell streams are packed by entry, so no cache
// line is fetched for a cell the tile does not evaluate: per T entry (owned
// then halo, tile by tile) its stress constants in five pairs, per U entry
// its stepu constants in seven pairs, and per owned T cell its 12 stresses in
// six pairs (a halo entry reads its owner's, hidx).  Velocities keep the
// block layout (the stencil gathers read neighbours by offset).
struct Fused2 {
  const Tile2 *tiles;
  const int32_t *tcell, *ucell;   // linear cells
  const uint16_t *tlds, *ulds;    // their LDS entries
  const int32_t *hidx;            // per T entry: packed stress index (its owner's)
  const double2 *tc, *uc;         // [5][et] and [7][eu] packed constants
  long et, eu, no;
  const double2 *sg_in;           // [6][no] stresses in pairs before the subcycle
  double2 *sg_out;                //          and after it
  double2 *vel_out;
};

// stress constants of T entry e: (dxt, dyt), (strength, dmin), (cxp, cyp),
// (cxm, cym), (dxhy, dyhx)
__global__ void k2_pack_t(View2 v, const int32_t *tcell, double2 *tc, long et) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= et) return;
  const long c = tcell[e];
  tc[e] = make_double2(v.dxt[c], v.dyt[c]);
  tc[et + e] = make_double2(v.strength[c], v.dmin[c]);
  tc[2 * et + e] = make_double2(v.cxp[c], v.cyp[c]);
  tc[3 * et + e] = make_double2(v.cxm[c], v.cym[c]);
  tc[4 * et + e] = make_double2(v.dxhy[c], v.dyhx[c]);
}
// stepu constants of U entry e: (cw, aix), (uocn, vocn), (waterx, watery),
// (forcex, forcey), (umassdti, fm), (uarear, tbu), (uvel_init, vvel_init)
__global__ void k2_pack_u(View2 v, const int32_t *ucell, double2 *uc, long eu) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= eu) return;
  const long c = ucell[e];
  uc[e] = make_double2(v.cw[c], v.aix[c]);
  uc[eu + e] = make_double2(v.uocn[c], v.vocn[c]);
  uc[2 * eu + e] = make_double2(v.waterx[c], v.watery[c]);
  uc[3 * eu + e] = make_double2(v.forcex[c], v.forcey[c]);
  uc[4 * eu + e] = make_double2(v.umassdti[c], v.fm[c]);
  uc[5 * eu + e] = make_double2(v.uarear[c], v.tbu[c]);
  uc[6 * eu + e] = v.vinit[c];
}
// the stresses of the owned T cells (ocell, packed order) into pairs, and back
__global__ void k2_pack_sig(const double *sig, long n, const int32_t *ocell, double2 *sg, long no) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= no) return;
  const long c = ocell[p];
#pragma unroll
  for (int q = 0; q < 6; ++q) sg[q * no + p] = make_double2(sig[2 * q * n + c], sig[(2 * q + 1) * n + c]);
}
__global__ void k2_unpack_sig(double *sig, long n, const int32_t *ocell, const double2 *sg, long no) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= no) return;
  const long c = ocell[p];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double2 s = sg[q * no + p];
    sig[2 * q * n + c] = s.x;
    sig[(2 * q + 1) * n + c] = s.y;
  }
}

// NT: the packed stresses (read and written once per launch) and str1..8 as
// non-temporal loads and stores (evp1d_cache.hpp), so that the last-level
// cache can keep the packed constants, tables and velocities between launches.
template <bool NT>
__global__ __launch_bounds__(kB2F) void k2_fused(View2 v, Fused2 f, int last) {
  __shared__ double2 lds[4 * kCap2];
  const Tile2 t = f.tiles[tile_of(blockIdx.x, gridDim.x, true)];   // each XCD a contiguous band of tiles
  for (int k = threadIdx.x; k < 4 * kCap2; k += kB2F) lds[k] = make_double2(0.0, 0.0);
  __syncthreads();
  const int nT = t.nto + t.nth;
  for (int k = threadIdx.x; k < nT; k += kB2F) {
    const long e = t.t0 + k;
    const long c = f.tcell[e];
    const long p = f.hidx[e];
    const int el = f.tlds[e];
    const double2 cc = v.vel[c];
    const double2 ve = v.vel[c - 1];
    const double2 vs = v.vel[c - v.nxb];
    const double2 vn = v.vel[c - v.nxb - 1];
    double sig[12];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const double2 s = in2<NT>(f.sg_in + (q * f.no + p));
      sig[2 * q] = s.x;
      sig[2 * q + 1] = s.y;
    }
    const double2 m0 = f.tc[e], m1 = f.tc[f.et + e], m2 = f.tc[2 * f.et + e], m3 = f.tc[3 * f.et + e],
                  m4 = f.tc[4 * f.et + e];
    evp1d::StressOut o;
    evp1d::stress_core(v.s, cc.x, cc.y, ve.x, ve.y, vs.x, vs.y, vn.x, vn.y, m0.x, m0.y, m1.x, m2.x, m2.y, m3.x,
                       m3.y, m1.y, m4.x, m4.y, sig, o);
    const double2 p15 = make_double2(o.str[0], o.str[4]), p27 = make_double2(o.str[1], o.str[6]);
    const double2 p36 = make_double2(o.str[2], o.str[5]), p48 = make_double2(o.str[3], o.str[7]);
    if (k < t.nto) {
      out2<NT>(f.sg_out + p, make_double2(o.sp[0], o.sp[1]));
      out2<NT>(f.sg_out + (f.no + p), make_double2(o.sp[2], o.sp[3]));
      out2<NT>(f.sg_out + (2 * f.no + p), make_double2(o.sm[0], o.sm[1]));
      out2<NT>(f.sg_out + (3 * f.no + p), make_double2(o.sm[2], o.sm[3]));
      out2<NT>(f.sg_out + (4 * f.no + p), make_double2(o.s12[0], o.s12[1]));
      out2<NT>(f.sg_out + (5 * f.no + p), make_double2(o.s12[2], o.s12[3]));
      if (last) {
        out2<NT>(v.s15 + c, p15);
        out2<NT>(v.s27 + c, p27);
        out2<NT>(v.s36 + c, p36);
        out2<NT>(v.s48 + c, p48);
      }
    }
    lds[el] = p15;
    lds[kCap2 + el] = p27;
    lds[2 * kCap2 + el] = p36;
    lds[3 * kCap2 + el] = p48;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < t.nu; k += kB2F) {
    const long e = t.u0 + k;
    const long c = f.ucell[e];
    const int el = f.ulds[e];
    const double2 a = lds[el];                        // (i, j): str1, str5
    const double2 b = lds[kCap2 + el + 1];            // (i+1, j): str2, str7
    const double2 g = lds[2 * kCap2 + el + kTW + 1];  // (i, j+1): str3, str6
    const double2 h = lds[3 * kCap2 + el + kTW + 2];  // (i+1, j+1): str4, str8
    const double2 vv = v.vel[c];
    const double2 k0 = f.uc[e], k1 = f.uc[f.eu + e], k2 = f.uc[2 * f.eu + e], k3 = f.uc[3 * f.eu + e];
    const double2 k4 = f.uc[4 * f.eu + e], k5 = f.uc[5 * f.eu + e], k6 = f.uc[6 * f.eu + e];
    evp1d::StepuOut o;
    evp1d::stepu_point(v.s, vv.x, vv.y, k0.x, k0.y, k1.x, k1.y, k2.x, k2.y, k3.x, k3.y, k4.x, k4.y, k5.x, k6.x,
                       k6.y, k5.y, a.x, b.x, g.x, h.x, a.y, g.y, b.y, h.y, o);
    f.vel_out[c] = make_double2(o.uvel, o.vvel);
    if (last) {
      v.strintx[c] = o.strintx;
      v.strinty[c] = o.strinty;
      v.taubx[c] = -(o.uvel * o.Cb);
      v.tauby[c] = -(o.vvel * o.Cb);
    }
  }
}

// halo update of uvel and vvel (vars.F90:113-142): the fill, then the copies
__global__ void k2_fill(double2 *vel, const int32_t *cells, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) vel[cells[k]] = make_double2(0.0, 0.0);
}
__global__ void k2_copy(double2 *vel, const int32_t *dst, const int32_t *src, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) vel[dst[k]] = src[k] >= 0 ? vel[src[k]] : make_double2(0.0, 0.0);
}

// one component of a double2 array from / to a plain array
__global__ void k2_put_comp(double2 *dst, const double *src, long n, int comp) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (comp == 0) dst[k].x = src[k];
  else dst[k].y = src[k];
}
__global__ void k2_get_comp(double *dst, const double2 *src, long n, int comp) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = comp == 0 ? src[k].x : src[k].y;
}

unsigned blocks2(long n) { return (unsigned)((n + kB2 - 1) / kB2); }

}  // namespace

struct evp2d_handle {
  int nxb = 0, nyb = 0, nblocks = 0;
  long plane = 0, n = 0;
  evp1d_params p{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void *> allocs;
  View2 v{};
  double *plain[P_COUNT] = {};
  double2 *str[4] = {};
  double *staging = nullptr;
  // cell lists per block (linear cells), concatenated on the device
  std::vector<std::vector<int32_t>> tl, ul;
  int32_t *tcell_dev = nullptr, *ucell_dev = nullptr;
  bool cells_dirty = true;   // device tables out of date
  bool lists_dirty = true;   // the T/U lists changed: str must read as zero again
  // halo (used when nblocks > 1 and set)
  bool halo_on = false;
  std::vector<int32_t> fill_h, dst_h, src_h;
  int32_t *fill_dev = nullptr, *dst_dev = nullptr, *src_dev = nullptr;
  // fused form (default): tile tables and the second velocity / stress buffers
  int mode = EVP2D_MODE_FUSED;
  int cache_hints = -1;   // k2_fused's cache hints: -1 chosen by size (hints2), 0 off, 1 on
  Tile2 *tiles_dev = nullptr;
  int32_t *ftcell_dev = nullptr, *fucell_dev = nullptr;
  uint16_t *ftlds_dev = nullptr, *fulds_dev = nullptr;
  int32_t *hidx_dev = nullptr, *ocell_dev = nullptr;   // per T entry: packed stress index; per index: cell
  double2 *tc_dev = nullptr, *uc_dev = nullptr;        // packed constants (per run)
  double2 *sg_dev[2] = {nullptr, nullptr};             // packed stresses during a run
  long et = 0, eu = 0, no = 0;                          // T entries, U entries, owned T cells
  int ntiles = 0;
  double2 *vel_alt = nullptr;
  // captured subcycle loop
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_nsub = -1;
  bool graph_dirty = true;
};

namespace {

void drop_graph2(evp2d_handle *h) {
  if (h->exec) (void)hipGraphExecDestroy(h->exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->exec = nullptr;
  h->graph = nullptr;
  h->graph_nsub = -1;
}

void free_lists2(evp2d_handle *h) {
  void *bufs[16] = {h->tcell_dev, h->ucell_dev, h->fill_dev, h->dst_dev, h->src_dev,   h->tiles_dev,
                    h->ftcell_dev, h->fucell_dev, h->ftlds_dev, h->fulds_dev, h->hidx_dev, h->ocell_dev,
                    h->tc_dev,    h->uc_dev,     h->sg_dev[0], h->sg_dev[1]};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  h->tcell_dev = h->ucell_dev = h->fill_dev = h->dst_dev = h->src_dev = nullptr;
  h->tiles_dev = nullptr;
  h->ftcell_dev = h->fucell_dev = nullptr;
  h->ftlds_dev = h->fulds_dev = nullptr;
  h->hidx_dev = h->ocell_dev = nullptr;
  h->tc_dev = h->uc_dev = h->sg_dev[0] = h->sg_dev[1] = nullptr;
  h->et = h->eu = h->no = 0;
  h->ntiles = 0;
}

void release2(evp2d_handle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph2(h);
  free_lists2(h);
  if (h->vel_alt) (void)hipFree(h->vel_alt);
  for (void *a : h->allocs) (void)hipFree(a);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int alloc2(evp2d_handle *h, void **ptr, size_t bytes) {
  const hipError_t e = hipMalloc(ptr, std::max<size_t>(bytes, 256));
  if (e != hipSuccess) return fail2(EVP1D_ENOMEM, "evp2d_create: hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  h->allocs.push_back(*ptr);
  HIP2(hipMemsetAsync(*ptr, 0, std::max<size_t>(bytes, 256), h->stream));
  return EVP1D_OK;
}

// The fused form's tables from the per-block lists: tile (tx, ty) of block
// b covers cells i in [tx*kTW + 1, tx*kTW + kTW], j in [ty*kTH + 1, ty*kTH +
// kTH].  A T cell is owned by its tile, and is a halo entry of the tile west
// of it when it is on its tile's first column, of the tile south of it when
// on the first row, and of the south-west tile when both; LDS entry of cell
// (i, j) in tile (tx, ty): (j-1 - ty*kTH) * (kTW+1) + (i-1 - tx*kTW).
int build_tiles2(evp2d_handle *h) {
  const int ntx = (h->nxb + kTW - 1) / kTW, nty = (h->nyb + kTH - 1) / kTH;
  const long nt_all = (long)h->nblocks * ntx * nty;
  std::vector<std::vector<int32_t>> own((size_t)nt_all), halo((size_t)nt_all), uc((size_t)nt_all);
  std::vector<std::vector<uint16_t>> own_e((size_t)nt_all), halo_e((size_t)nt_all), uc_e((size_t)nt_all);
  auto tile = [&](int b, int tx, int ty) { return ((long)b * nty + ty) * ntx + tx; };
  auto entry = [&](int i, int j, int tx, int ty) {
    const int di = i - 1 - tx * kTW, dj = j - 1 - ty * kTH;
    return (uint16_t)(dj * (kTW + 1) + di);
  };
  for (int b = 0; b < h->nblocks; ++b) {
    const long base = (long)b * h->plane;
    for (int32_t c : h->tl[(size_t)b]) {
      const long l = c - base;
      const int i = (int)(l % h->nxb) + 1, j = (int)(l / h->nxb) + 1;
      const int tx = (i - 1) / kTW, ty = (j - 1) / kTH;
      own[(size_t)tile(b, tx, ty)].push_back(c);
      own_e[(size_t)tile(b, tx, ty)].push_back(entry(i, j, tx, ty));
      const bool col0 = (i - 1) % kTW == 0 && tx > 0, row0 = (j - 1) % kTH == 0 && ty > 0;
      if (col0) { halo[(size_t)tile(b, tx - 1, ty)].push_back(c); halo_e[(size_t)tile(b, tx - 1, ty)].push_back(entry(i, j, tx - 1, ty)); }
      if (row0) { halo[(size_t)tile(b, tx, ty - 1)].push_back(c); halo_e[(size_t)tile(b, tx, ty - 1)].push_back(entry(i, j, tx, ty - 1)); }
      if (col0 && row0) {
        halo[(size_t)tile(b, tx - 1, ty - 1)].push_back(c);
        halo_e[(size_t)tile(b, tx - 1, ty - 1)].push_back(entry(i, j, tx - 1, ty - 1));
      }
    }
    for (int32_t c : h->ul[(size_t)b]) {
      const long l = c - base;
      const int i = (int)(l % h->nxb) + 1, j = (int)(l / h->nxb) + 1;
      const int tx = (i - 1) / kTW, ty = (j - 1) / kTH;
      uc[(size_t)tile(b, tx, ty)].push_back(c);
      uc_e[(size_t)tile(b, tx, ty)].push_back(entry(i, j, tx, ty));
    }
  }
  std::vector<Tile2> tiles;
  std::vector<int32_t> tcell, ucell, ocell;
  std::vector<uint16_t> tlds, ulds;
  // packed stress index of every owned T cell: its position among the owned
  // entries, tile by tile
  std::vector<int32_t> pidx((size_t)h->n, -1);
  for (long t = 0; t < nt_all; ++t)
    for (int32_t c : own[(size_t)t]) {
      pidx[(size_t)c] = (int32_t)ocell.size();
      ocell.push_back(c);
    }
  std::vector<int32_t> hidx;
  for (long t = 0; t < nt_all; ++t) {
    const size_t k = (size_t)t;
    if (own[k].empty() && uc[k].empty()) continue;
    for (int32_t c : own[k]) hidx.push_back(pidx[(size_t)c]);
    for (int32_t c : halo[k]) hidx.push_back(pidx[(size_t)c]);
    Tile2 d;
    d.t0 = (int32_t)tcell.size();
    d.nto = (int32_t)own[k].size();
    d.nth = (int32_t)halo[k].size();
    d.u0 = (int32_t)ucell.size();
    d.nu = (int32_t)uc[k].size();
    tcell.insert(tcell.end(), own[k].begin(), own[k].end());
    tcell.insert(tcell.end(), halo[k].begin(), halo[k].end());
    tlds.insert(tlds.end(), own_e[k].begin(), own_e[k].end());
    tlds.insert(tlds.end(), halo_e[k].begin(), halo_e[k].end());
    ucell.insert(ucell.end(), uc[k].begin(), uc[k].end());
    ulds.insert(ulds.end(), uc_e[k].begin(), uc_e[k].end());
    tiles.push_back(d);
  }
  // every entry inside its tile's table, every stepu read too (the kernel
  // trusts these): the U entry e reads e+1, e+kTW+1 and e+kTW+2
  for (uint16_t e : tlds)
    if (e >= kCap2) return fail2(EVP1D_EINDEX, "evp2d: T entry %d outside the tile table", (int)e);
  for (const Tile2 &d : tiles)
    for (int k = 0; k < d.nu; ++k) {
      const int e = ulds[(size_t)d.u0 + k];
      if (e % (kTW + 1) >= kTW || e / (kTW + 1) >= kTH)
        return fail2(EVP1D_EINDEX, "evp2d: U entry %d outside its tile", e);
    }
  auto up32 = [&](int32_t **dst, const std::vector<int32_t> &src) -> int {
    HIP2(hipMalloc((void **)dst, std::max<size_t>(src.size(), 1) * sizeof(int32_t)));
    if (!src.empty()) HIP2(hipMemcpyAsync(*dst, src.data(), src.size() * 4, hipMemcpyHostToDevice, h->stream));
    return EVP1D_OK;
  };
  auto up16 = [&](uint16_t **dst, const std::vector<uint16_t> &src) -> int {
    HIP2(hipMalloc((void **)dst, std::max<size_t>(src.size(), 1) * sizeof(uint16_t)));
    if (!src.empty()) HIP2(hipMemcpyAsync(*dst, src.data(), src.size() * 2, hipMemcpyHostToDevice, h->stream));
    return EVP1D_OK;
  };
  HIP2(hipMalloc((void **)&h->tiles_dev, std::max<size_t>(tiles.size(), 1) * sizeof(Tile2)));
  if (!tiles.empty())
    HIP2(hipMemcpyAsync(h->tiles_dev, tiles.data(), tiles.size() * sizeof(Tile2), hipMemcpyHostToDevice, h->stream));
  RC2(up32(&h->ftcell_dev, tcell));
  RC2(up32(&h->fucell_dev, ucell));
  RC2(up16(&h->ftlds_dev, tlds));
  RC2(up16(&h->fulds_dev, ulds));
  for (int32_t p : hidx)   // every halo T cell is some tile's owned cell
    if (p < 0) return fail2(EVP1D_EINDEX, "evp2d: a halo T entry has no owning tile");
  RC2(up32(&h->hidx_dev, hidx));
  RC2(up32(&h->ocell_dev, ocell));
  h->et = (long)tcell.size();
  h->eu = (long)ucell.size();
  h->no = (long)ocell.size();
  HIP2(hipMalloc((void **)&h->tc_dev, std::max<size_t>((size_t)5 * h->et, 1) * sizeof(double2)));
  HIP2(hipMalloc((void **)&h->uc_dev, std::max<size_t>((size_t)7 * h->eu, 1) * sizeof(double2)));
  for (int b = 0; b < 2; ++b)
    HIP2(hipMalloc((void **)&h->sg_dev[b], std::max<size_t>((size_t)6 * h->no, 1) * sizeof(double2)));
  HIP2(hipStreamSynchronize(h->stream));   // the host vectors go out of scope
  h->ntiles = (int)tiles.size();
  return EVP1D_OK;
}

// upload the concatenated lists (and the halo tables) when they changed;
// a host allocation failure while building them returns EVP1D_ENOMEM
int ensure2_impl(evp2d_handle *h);
int ensure2(evp2d_handle *h) {
  if (!h->cells_dirty) return EVP1D_OK;
  return evp1d::no_throw("evp2d lists", [h] { return ensure2_impl(h); });
}
int ensure2_impl(evp2d_handle *h) {
  HIP2(hipStreamSynchronize(h->stream));
  drop_graph2(h);
  free_lists2(h);
  std::vector<int32_t> t, u;
  for (int b = 0; b < h->nblocks; ++b) {
    t.insert(t.end(), h->tl[(size_t)b].begin(), h->tl[(size_t)b].end());
    u.insert(u.end(), h->ul[(size_t)b].begin(), h->ul[(size_t)b].end());
  }
  auto up = [&](int32_t **dst, const std::vector<int32_t> &src) -> int {
    HIP2(hipMalloc((void **)dst, std::max<size_t>(src.size(), 1) * sizeof(int32_t)));
    if (!src.empty())
      HIP2(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    return EVP1D_OK;
  };
  RC2(up(&h->tcell_dev, t));
  RC2(up(&h->ucell_dev, u));
  RC2(up(&h->fill_dev, h->fill_h));
  RC2(up(&h->dst_dev, h->dst_h));
  RC2(up(&h->src_dev, h->src_h));
  RC2(build_tiles2(h));
  // str outside the lists must read as zero (the reference clears it)
  if (h->lists_dirty)
    for (double2 *s : h->str) HIP2(hipMemsetAsync(s, 0, (size_t)h->n * sizeof(double2), h->stream));
  HIP2(hipStreamSynchronize(h->stream));
  h->lists_dirty = false;
  h->v.tcell = h->tcell_dev;
  h->v.ucell = h->ucell_dev;
  h->v.nt = (int)t.size();
  h->v.nu = (int)u.size();
  h->cells_dirty = false;
  h->graph_dirty = true;
  return EVP1D_OK;
}

int launch_stress2(evp2d_handle *h) {
  if (h->v.nt > 0) hipLaunchKernelGGL(k2_stress, dim3(blocks2(h->v.nt)), dim3(kB2), 0, h->stream, h->v);
  HIP2(hipGetLastError());
  return EVP1D_OK;
}
int launch_stepu2(evp2d_handle *h) {
  if (h->v.nu > 0) hipLaunchKernelGGL(k2_stepu, dim3(blocks2(h->v.nu)), dim3(kB2), 0, h->stream, h->v);
  HIP2(hipGetLastError());
  return EVP1D_OK;
}
int launch_halo2(evp2d_handle *h) {
  if (!h->halo_on || h->nblocks <= 1) return EVP1D_OK;   // ice_haloUpdate returns at max_blocks == 1
  const int nf = (int)h->fill_h.size(), nc = (int)h->dst_h.size();
  if (nf > 0) hipLaunchKernelGGL(k2_fill, dim3(blocks2(nf)), dim3(kB2), 0, h->stream, h->v.vel, h->fill_dev, nf);
  if (nc > 0)
    hipLaunchKernelGGL(k2_copy, dim3(blocks2(nc)), dim3(kB2), 0, h->stream, h->v.vel, h->dst_dev, h->src_dev, nc);
  HIP2(hipGetLastError());
  return EVP1D_OK;
}
int issue2(evp2d_handle *h, int nsub) {
  for (int k = 0; k < nsub; ++k) {
    RC2(launch_stress2(h));
    RC2(launch_stepu2(h));
    RC2(launch_halo2(h));
  }
  return EVP1D_OK;
}

// halo update on the buffer a subcycle wrote (out)
int launch_halo2_on(evp2d_handle *h, double2 *vel) {
  if (!h->halo_on || h->nblocks <= 1) return EVP1D_OK;
  const int nf = (int)h->fill_h.size(), nc = (int)h->dst_h.size();
  if (nf > 0) hipLaunchKernelGGL(k2_fill, dim3(blocks2(nf)), dim3(kB2), 0, h->stream, vel, h->fill_dev, nf);
  if (nc > 0) hipLaunchKernelGGL(k2_copy, dim3(blocks2(nc)), dim3(kB2), 0, h->stream, vel, h->dst_dev, h->src_dev, nc);
  HIP2(hipGetLastError());
  return EVP1D_OK;
}

// Whether k2_fused uses the cache hints.  By default (hints_pay): what every
// launch reads again is the packed constants (5 pairs per T entry, 7 per U
// entry), the tables (20 B per T entry, 12 per U entry) and the two velocity
// arrays (16 B per cell each); the packed stresses add 192 B per owned T cell.
bool hints2(const evp2d_handle *h) {
  if (h->cache_hints >= 0) return h->cache_hints != 0;
  const double reread = 100.0 * h->et + 124.0 * h->eu + 32.0 * h->n;
  return hints_pay(reread, reread + 192.0 * h->no);
}

// nsub fused subcycles.  The velocities start in the handle's buffer (X),
// are copied to the second one (Y), and the subcycles alternate so that the
// last one writes X: subcycle k writes X when nsub - k is even.  The
// constants and stresses are packed by entry first (k2_pack_*); the
// stresses alternate between two packed buffers and are put back into the
// handle's block arrays after the last subcycle.
int issue_fused2(evp2d_handle *h, int nsub) {
  const View2 X = h->v;
  View2 Y = h->v;
  Y.vel = h->vel_alt;
  HIP2(hipMemcpyAsync(Y.vel, X.vel, (size_t)h->n * sizeof(double2), hipMemcpyDeviceToDevice, h->stream));
  if (h->et > 0) hipLaunchKernelGGL(k2_pack_t, dim3(blocks2(h->et)), dim3(kB2), 0, h->stream, X, h->ftcell_dev, h->tc_dev, h->et);
  if (h->eu > 0) hipLaunchKernelGGL(k2_pack_u, dim3(blocks2(h->eu)), dim3(kB2), 0, h->stream, X, h->fucell_dev, h->uc_dev, h->eu);
  if (h->no > 0)
    hipLaunchKernelGGL(k2_pack_sig, dim3(blocks2(h->no)), dim3(kB2), 0, h->stream, (const double *)X.sig, h->n,
                       (const int32_t *)h->ocell_dev, h->sg_dev[0], h->no);
  HIP2(hipGetLastError());
  const auto fused = hints2(h) ? k2_fused<true> : k2_fused<false>;
  int cur = 0;
  for (int k = 1; k <= nsub; ++k) {
    const bool to_x = ((nsub - k) % 2) == 0;
    const View2 &in = to_x ? Y : X;
    const View2 &out = to_x ? X : Y;
    Fused2 f{h->tiles_dev, h->ftcell_dev, h->fucell_dev, h->ftlds_dev, h->fulds_dev, h->hidx_dev,
             h->tc_dev,    h->uc_dev,     h->et,         h->eu,        h->no,         h->sg_dev[cur],
             h->sg_dev[1 - cur], out.vel};
    if (h->ntiles > 0)
      hipLaunchKernelGGL(fused, dim3(h->ntiles), dim3(kB2F), 0, h->stream, in, f, k == nsub ? 1 : 0);
    HIP2(hipGetLastError());
    RC2(launch_halo2_on(h, out.vel));
    cur = 1 - cur;
  }
  if (h->no > 0)
    hipLaunchKernelGGL(k2_unpack_sig, dim3(blocks2(h->no)), dim3(kB2), 0, h->stream, X.sig, h->n,
                       (const int32_t *)h->ocell_dev, (const double2 *)h->sg_dev[cur], h->no);
  HIP2(hipGetLastError());
  return EVP1D_OK;
}

// field id -> where it lives
struct Loc2 { int kind; int idx; };   // kind 0 plain, 1 vel comp, 2 vinit comp, 3 sig, 4 str comp
bool field_loc(int32_t f, Loc2 &l) {
  static const int plain_of[EVP2D_STRESSP_1] = {
      P_STRENGTH, -1, -1, P_DXT, P_DYT, P_DXHY, P_DYHX, P_CXP, P_CYP, P_CXM, P_CYM, P_DMIN, P_UAREAR,
      P_CW, P_AIX, P_UOCN, P_VOCN, P_WATERX, P_WATERY, P_FORCEX, P_FORCEY, P_UMASSDTI, P_FM, P_STRINTX,
      P_STRINTY, P_TBU};
  if (f < 0 || f >= EVP2D_NFIELDS) return false;
  if (f == EVP2D_UVEL || f == EVP2D_VVEL) { l = {1, f == EVP2D_UVEL ? 0 : 1}; return true; }
  if (f < EVP2D_STRESSP_1) { l = {0, plain_of[f]}; return true; }
  if (f < EVP2D_UVEL_INIT) { l = {3, f - EVP2D_STRESSP_1}; return true; }
  if (f == EVP2D_UVEL_INIT || f == EVP2D_VVEL_INIT) { l = {2, f == EVP2D_UVEL_INIT ? 0 : 1}; return true; }
  if (f == EVP2D_TAUBX) { l = {0, P_TAUBX}; return true; }
  if (f == EVP2D_TAUBY) { l = {0, P_TAUBY}; return true; }
  // str1..8 -> pair (0: 1/5, 1: 2/7, 2: 3/6, 3: 4/8) * 2 + component
  static const int pr[8] = {0, 2, 4, 6, 1, 5, 3, 7};
  l = {4, pr[f - EVP2D_STR1]};
  return true;
}

}  // namespace

extern "C" {

int evp2d_create(int32_t nx_block, int32_t ny_block, int32_t max_blocks, const evp1d_params *p,
                 evp2d_handle **out) {
  if (!out) return fail2(EVP1D_EINVAL, "evp2d_create: out is NULL");
  *out = nullptr;
  if (!p) return fail2(EVP1D_EINVAL, "evp2d_create: params is NULL");
  if (nx_block < 2 || ny_block < 2 || max_blocks < 1)
    return fail2(EVP1D_EINVAL, "evp2d_create: need nx_block, ny_block >= 2 and max_blocks >= 1 (%d, %d, %d)",
                 nx_block, ny_block, max_blocks);
  const long plane = (long)nx_block * ny_block, n = plane * max_blocks;
  if (n > 0x7fffffffL) return fail2(EVP1D_EINVAL, "evp2d_create: %ld cells exceed int32", n);
  evp2d_handle *h = new (std::nothrow) evp2d_handle();
  if (!h) return fail2(EVP1D_ENOMEM, "evp2d_create: host allocation failed");
  h->nxb = nx_block;
  h->nyb = ny_block;
  h->nblocks = max_blocks;
  h->plane = plane;
  h->n = n;
  h->p = *p;
  h->tl.assign((size_t)max_blocks, {});
  h->ul.assign((size_t)max_blocks, {});
  int rc = EVP1D_OK;
  auto chk = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && !rc) rc = fail2(EVP1D_EHIP, "%s: %s", what, hipGetErrorString(e));
  };
  chk(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
  chk(hipEventCreate(&h->ev0), "hipEventCreate");
  chk(hipEventCreate(&h->ev1), "hipEventCreate");
  const size_t nD = (size_t)n * sizeof(double), nP = (size_t)n * sizeof(double2);
  View2 &v = h->v;
  auto A = [&](void **ptr, size_t bytes) { if (!rc) rc = alloc2(h, ptr, bytes); };
  A((void **)&v.vel, nP);
  A((void **)&v.vinit, nP);
  A((void **)&v.sig, 12 * nD);
  for (auto &s : h->str) A((void **)&s, nP);
  for (auto &q : h->plain) A((void **)&q, nD);
  A((void **)&h->staging, nD);
  if (!rc) rc = hipStreamSynchronize(h->stream) == hipSuccess ? EVP1D_OK : fail2(EVP1D_EHIP, "evp2d_create: sync");
  if (rc) { release2(h); return rc; }
  v.n = n;
  v.nxb = nx_block;
  v.s15 = h->str[0]; v.s27 = h->str[1]; v.s36 = h->str[2]; v.s48 = h->str[3];
  double **P = h->plain;
  v.strength = P[P_STRENGTH]; v.dxt = P[P_DXT]; v.dyt = P[P_DYT]; v.dxhy = P[P_DXHY]; v.dyhx = P[P_DYHX];
  v.cxp = P[P_CXP]; v.cyp = P[P_CYP]; v.cxm = P[P_CXM]; v.cym = P[P_CYM]; v.dmin = P[P_DMIN];
  v.uarear = P[P_UAREAR]; v.cw = P[P_CW]; v.aix = P[P_AIX]; v.uocn = P[P_UOCN]; v.vocn = P[P_VOCN];
  v.waterx = P[P_WATERX]; v.watery = P[P_WATERY]; v.forcex = P[P_FORCEX]; v.forcey = P[P_FORCEY];
  v.umassdti = P[P_UMASSDTI]; v.fm = P[P_FM]; v.strintx = P[P_STRINTX]; v.strinty = P[P_STRINTY];
  v.tbu = P[P_TBU]; v.taubx = P[P_TAUBX]; v.tauby = P[P_TAUBY];
  v.s = scalars2(*p);
  *out = h;
  return EVP1D_OK;
}

int evp2d_destroy(evp2d_handle *h) {
  release2(h);
  return EVP1D_OK;
}

int evp2d_set_params(evp2d_handle *h, const evp1d_params *p) {
  if (!h || !p) return fail2(EVP1D_EINVAL, "evp2d_set_params: NULL argument");
  h->p = *p;
  h->v.s = scalars2(*p);
  h->graph_dirty = true;
  return EVP1D_OK;
}

int evp2d_set_cells(evp2d_handle *h, int32_t iblk, int32_t icellt, const int32_t *indxti, const int32_t *indxtj,
                    int32_t icellu, const int32_t *indxui, const int32_t *indxuj) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_set_cells: NULL handle");
  if (iblk < 1 || iblk > h->nblocks) return fail2(EVP1D_EINVAL, "evp2d_set_cells: block %d outside 1..%d", iblk, h->nblocks);
  const long cells = h->plane;
  if (icellt < 0 || icellt > cells || icellu < 0 || icellu > cells)
    return fail2(EVP1D_EINVAL, "evp2d_set_cells: icellt=%d icellu=%d outside 0..%ld", icellt, icellu, cells);
  if ((icellt > 0 && (!indxti || !indxtj)) || (icellu > 0 && (!indxui || !indxuj)))
    return fail2(EVP1D_EINVAL, "evp2d_set_cells: NULL list");
  const long base = (long)(iblk - 1) * h->plane;
  std::vector<char> seen((size_t)cells, 0);
  std::vector<int32_t> t((size_t)icellt), u((size_t)icellu);
  // stress reads (i-1, j-1) .. (i, j): 2 <= i <= nx_block, 2 <= j <= ny_block
  for (int k = 0; k < icellt; ++k) {
    const int i = indxti[k], j = indxtj[k];
    if (i < 2 || i > h->nxb || j < 2 || j > h->nyb)
      return fail2(EVP1D_EINDEX, "evp2d_set_cells: T cell %d = (%d, %d) needs 2 <= i <= %d, 2 <= j <= %d", k + 1, i,
                   j, h->nxb, h->nyb);
    const long c = (long)(j - 1) * h->nxb + (i - 1);
    if (seen[(size_t)c]++) return fail2(EVP1D_EINDEX, "evp2d_set_cells: T cell (%d, %d) listed twice", i, j);
    t[(size_t)k] = (int32_t)(base + c);
  }
  std::fill(seen.begin(), seen.end(), 0);
  // stepu reads (i, j) .. (i+1, j+1): 1 <= i < nx_block, 1 <= j < ny_block
  for (int k = 0; k < icellu; ++k) {
    const int i = indxui[k], j = indxuj[k];
    if (i < 1 || i >= h->nxb || j < 1 || j >= h->nyb)
      return fail2(EVP1D_EINDEX, "evp2d_set_cells: U cell %d = (%d, %d) needs 1 <= i < %d, 1 <= j < %d", k + 1, i, j,
                   h->nxb, h->nyb);
    const long c = (long)(j - 1) * h->nxb + (i - 1);
    if (seen[(size_t)c]++) return fail2(EVP1D_EINDEX, "evp2d_set_cells: U cell (%d, %d) listed twice", i, j);
    u[(size_t)k] = (int32_t)(base + c);
  }
  h->tl[(size_t)(iblk - 1)] = std::move(t);
  h->ul[(size_t)(iblk - 1)] = std::move(u);
  h->cells_dirty = true;
  h->lists_dirty = true;
  return EVP1D_OK;
}

int evp2d_set_halo(evp2d_handle *h, int32_t nblocks, const int32_t *ilo, const int32_t *jlo, const int32_t *ihi,
                   const int32_t *jhi, int32_t ncopies, const int32_t *src, const int32_t *dst) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_set_halo: NULL handle");
  if (nblocks < 0 || nblocks > h->nblocks || ncopies < 0)
    return fail2(EVP1D_EINVAL, "evp2d_set_halo: nblocks=%d (max %d), ncopies=%d", nblocks, h->nblocks, ncopies);
  if ((nblocks > 0 && (!ilo || !jlo || !ihi || !jhi)) || (ncopies > 0 && (!src || !dst)))
    return fail2(EVP1D_EINVAL, "evp2d_set_halo: NULL array");
  auto lin = [&](int i, int j, int b) { return (long)(b - 1) * h->plane + (long)(j - 1) * h->nxb + (i - 1); };
  // fill of the one-cell halo rows and columns of each local block (vars.F90:113-123)
  std::vector<char> mark((size_t)h->n, 0);
  std::vector<int32_t> fill;
  for (int b = 1; b <= nblocks; ++b) {
    const int il = ilo[b - 1], jl = jlo[b - 1], ih = ihi[b - 1], jh = jhi[b - 1];
    if (il < 2 || ih > h->nxb - 1 || jl < 2 || jh > h->nyb - 1 || il > ih + 1 || jl > jh + 1)
      return fail2(EVP1D_EINDEX, "evp2d_set_halo: block %d bounds (%d:%d, %d:%d) leave no halo inside %d x %d", b, il,
                   ih, jl, jh, h->nxb, h->nyb);
    auto add = [&](int i, int j) {
      const long c = lin(i, j, b);
      if (!mark[(size_t)c]) { mark[(size_t)c] = 1; fill.push_back((int32_t)c); }
    };
    for (int i = 1; i <= h->nxb; ++i) { add(i, jl - 1); add(i, jh + 1); }
    for (int j = 1; j <= h->nyb; ++j) { add(il - 1, j); add(ih + 1, j); }
  }
  // copies in list order (vars.F90:124-142); the last copy into a cell wins
  std::vector<long> last((size_t)h->n, -1);
  std::vector<int32_t> dl, sl;
  auto inside = [&](int i, int j, int b) { return i >= 1 && i <= h->nxb && j >= 1 && j <= h->nyb && b >= 1 && b <= h->nblocks; };
  for (int m = 0; m < ncopies; ++m) {
    const int is = src[3 * m], js = src[3 * m + 1], bs = src[3 * m + 2];
    const int id = dst[3 * m], jd = dst[3 * m + 1], bd = dst[3 * m + 2];
    if (bs > 0) {
      if (bd < 0) return fail2(EVP1D_EINVAL, "evp2d_set_halo: copy %d into a tripole buffer (dst block %d)", m + 1, bd);
      if (bd == 0) continue;   // the reference does nothing then
      if (!inside(is, js, bs) || !inside(id, jd, bd))
        return fail2(EVP1D_EINDEX, "evp2d_set_halo: copy %d (%d,%d,%d) -> (%d,%d,%d) outside the arrays", m + 1, is, js,
                     bs, id, jd, bd);
    } else if (bs == 0) {
      if (!inside(id, jd, bd))
        return fail2(EVP1D_EINDEX, "evp2d_set_halo: fill %d -> (%d,%d,%d) outside the arrays", m + 1, id, jd, bd);
    } else {
      continue;   // negative source block: no branch of the reference applies
    }
    const long c = lin(id, jd, bd);
    if (last[(size_t)c] >= 0) {   // an earlier copy into the same cell is overwritten
      dl[(size_t)last[(size_t)c]] = -1;
    }
    last[(size_t)c] = (long)dl.size();
    dl.push_back((int32_t)c);
    sl.push_back(bs > 0 ? (int32_t)lin(is, js, bs) : -1);
  }
  // drop overwritten copies; a copy that reads a cell some copy writes would
  // depend on the order of the list: not supported
  std::vector<int32_t> d2, s2;
  std::vector<char> isdst((size_t)h->n, 0);
  for (size_t k = 0; k < dl.size(); ++k)
    if (dl[k] >= 0) { d2.push_back(dl[k]); s2.push_back(sl[k]); isdst[(size_t)dl[k]] = 1; }
  for (size_t k = 0; k < s2.size(); ++k)
    if (s2[k] >= 0 && isdst[(size_t)s2[k]])
      return fail2(EVP1D_EINVAL, "evp2d_set_halo: copy %zu reads a cell another copy writes (chained copies)", k + 1);
  HIP2(hipStreamSynchronize(h->stream));
  h->fill_h = std::move(fill);
  h->dst_h = std::move(d2);
  h->src_h = std::move(s2);
  h->halo_on = nblocks > 0 || !h->dst_h.empty();
  h->cells_dirty = true;   // re-upload the tables
  return EVP1D_OK;
}

int evp2d_set_field(evp2d_handle *h, int32_t field, const double *a, int64_t n) {
  if (!h || !a) return fail2(EVP1D_EINVAL, "evp2d_set_field: NULL argument");
  Loc2 l;
  if (!field_loc(field, l)) return fail2(EVP1D_EINVAL, "evp2d_set_field: field %d unknown", field);
  if (l.kind == 4)
    return fail2(EVP1D_EINVAL, "evp2d_set_field: str1..8 are scratch, cleared by every stress call (read them with "
                               "evp2d_get_field)");
  if (n != h->n) return fail2(EVP1D_EINVAL, "evp2d_set_field: %lld values, need nx_block*ny_block*max_blocks = %ld",
                              (long long)n, h->n);
  const size_t bytes = (size_t)n * sizeof(double);
  HIP2(hipStreamSynchronize(h->stream));
  if (l.kind == 0 || l.kind == 3) {
    double *d = l.kind == 0 ? h->plain[l.idx] : h->v.sig + (size_t)l.idx * h->n;
    HIP2(hipMemcpyAsync(d, a, bytes, hipMemcpyHostToDevice, h->stream));
  } else {
    HIP2(hipMemcpyAsync(h->staging, a, bytes, hipMemcpyHostToDevice, h->stream));
    double2 *d = l.kind == 1 ? h->v.vel : const_cast<double2 *>(h->v.vinit);
    hipLaunchKernelGGL(k2_put_comp, dim3(blocks2(n)), dim3(kB2), 0, h->stream, d, (const double *)h->staging, (long)n,
                       l.idx);
    HIP2(hipGetLastError());
  }
  HIP2(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

int evp2d_get_field(evp2d_handle *h, int32_t field, double *a, int64_t n) {
  if (!h || !a) return fail2(EVP1D_EINVAL, "evp2d_get_field: NULL argument");
  Loc2 l;
  if (!field_loc(field, l)) return fail2(EVP1D_EINVAL, "evp2d_get_field: field %d unknown", field);
  if (n != h->n) return fail2(EVP1D_EINVAL, "evp2d_get_field: %lld values, need nx_block*ny_block*max_blocks = %ld",
                              (long long)n, h->n);
  RC2(ensure2(h));   // str reads as zero outside the lists once they are set
  const size_t bytes = (size_t)n * sizeof(double);
  if (l.kind == 0 || l.kind == 3) {
    const double *s = l.kind == 0 ? h->plain[l.idx] : h->v.sig + (size_t)l.idx * h->n;
    HIP2(hipMemcpyAsync(a, s, bytes, hipMemcpyDeviceToHost, h->stream));
  } else {
    const double2 *s = l.kind == 1 ? h->v.vel : (l.kind == 2 ? h->v.vinit : h->str[l.idx / 2]);
    const int comp = l.kind == 4 ? l.idx % 2 : l.idx;
    hipLaunchKernelGGL(k2_get_comp, dim3(blocks2(n)), dim3(kB2), 0, h->stream, h->staging, s, (long)n, comp);
    HIP2(hipGetLastError());
    HIP2(hipMemcpyAsync(a, h->staging, bytes, hipMemcpyDeviceToHost, h->stream));
  }
  HIP2(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

int evp2d_run(evp2d_handle *h, int32_t nsub) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_run: NULL handle");
  if (nsub < 0) return fail2(EVP1D_EINVAL, "evp2d_run: nsub=%d < 0", nsub);
  if (nsub == 0) return EVP1D_OK;
  RC2(ensure2(h));
  if (h->graph_dirty || h->graph_nsub != nsub || !h->exec) {
    drop_graph2(h);
    const bool fused = h->mode == EVP2D_MODE_FUSED;
    if (fused && !h->vel_alt) HIP2(hipMalloc((void **)&h->vel_alt, (size_t)h->n * sizeof(double2)));
    HIP2(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = fused ? issue_fused2(h, nsub) : issue2(h, nsub);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail2(EVP1D_EHIP, "evp2d_run: hipStreamEndCapture: %s", hipGetErrorString(e));
    h->graph = g;
    HIP2(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
    h->graph_nsub = nsub;
    h->graph_dirty = false;
  }
  HIP2(hipGraphLaunch(h->exec, h->stream));
  return EVP1D_OK;
}

int evp2d_set_mode(evp2d_handle *h, int32_t mode) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_set_mode: NULL handle");
  if (mode != EVP2D_MODE_PLAIN && mode != EVP2D_MODE_FUSED) return fail2(EVP1D_EINVAL, "evp2d_set_mode: mode %d", mode);
  if (mode != h->mode) {
    h->mode = mode;
    h->graph_dirty = true;
  }
  return EVP1D_OK;
}

int evp2d_get_mode(evp2d_handle *h, int32_t *mode) {
  if (!h || !mode) return fail2(EVP1D_EINVAL, "evp2d_get_mode: NULL argument");
  *mode = h->mode;
  return EVP1D_OK;
}

int evp2d_set_cache_hints(evp2d_handle *h, int32_t hints) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_set_cache_hints: NULL handle");
  if (hints < -1 || hints > 1) return fail2(EVP1D_EINVAL, "evp2d_set_cache_hints: hints=%d not -1, 0 or 1", hints);
  if (hints != h->cache_hints) {
    h->cache_hints = hints;
    h->graph_dirty = true;   // the captured launches name the kernel
  }
  return EVP1D_OK;
}

int evp2d_get_cache_hints(evp2d_handle *h, int32_t *hints, int32_t *active) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_get_cache_hints: NULL handle");
  RC2(ensure2(h));
  if (hints) *hints = h->cache_hints;
  if (active) *active = h->mode == EVP2D_MODE_FUSED && h->ntiles > 0 && hints2(h) ? 1 : 0;
  return EVP1D_OK;
}

int evp2d_stress(evp2d_handle *h) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_stress: NULL handle");
  RC2(ensure2(h));
  return launch_stress2(h);
}
int evp2d_stepu(evp2d_handle *h) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_stepu: NULL handle");
  RC2(ensure2(h));
  return launch_stepu2(h);
}
int evp2d_halo(evp2d_handle *h) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_halo: NULL handle");
  RC2(ensure2(h));
  return launch_halo2(h);
}

int evp2d_run_individual(evp2d_handle *h, int32_t which, int32_t nsub) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_run_individual: NULL handle");
  if (which < 0 || which > 2) return fail2(EVP1D_EINVAL, "evp2d_run_individual: which=%d", which);
  RC2(ensure2(h));
  if (which == 2) {   // fused subcycles launched one by one (no graph)
    if (nsub <= 0) return EVP1D_OK;
    if (!h->vel_alt) HIP2(hipMalloc((void **)&h->vel_alt, (size_t)h->n * sizeof(double2)));
    return issue_fused2(h, nsub);
  }
  for (int k = 0; k < nsub; ++k) RC2(which == 0 ? launch_stress2(h) : launch_stepu2(h));
  return EVP1D_OK;
}

int evp2d_synchronize(evp2d_handle *h) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_synchronize: NULL handle");
  HIP2(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

int evp2d_timer_start(evp2d_handle *h) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_timer_start: NULL handle");
  HIP2(hipEventRecord(h->ev0, h->stream));
  return EVP1D_OK;
}
int evp2d_timer_stop(evp2d_handle *h, float *ms) {
  if (!h || !ms) return fail2(EVP1D_EINVAL, "evp2d_timer_stop: NULL argument");
  HIP2(hipEventRecord(h->ev1, h->stream));
  HIP2(hipEventSynchronize(h->ev1));
  HIP2(hipEventElapsedTime(ms, h->ev0, h->ev1));
  return EVP1D_OK;
}

int evp2d_get_cells(evp2d_handle *h, int32_t *nt, int32_t *nu) {
  if (!h) return fail2(EVP1D_EINVAL, "evp2d_get_cells: NULL handle");
  long t = 0, u = 0;
  for (int b = 0; b < h->nblocks; ++b) {
    t += (long)h->tl[(size_t)b].size();
    u += (long)h->ul[(size_t)b].size();
  }
  if (nt) *nt = (int32_t)t;
  if (nu) *nu = (int32_t)u;
  return EVP1D_OK;
}

}  // extern "C"
