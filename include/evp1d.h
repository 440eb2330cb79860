/*
 * evp1d.h - C ABI of the MI355X EVP1D subcycle engine (libevp1d.so).
 *
 * This is the boundary that replaces the reference's kernel module `bench`
 * (src/bench_v2b.F90:62-312 `stress`, :444-549 `stepu`) and the subcycle loop
 * of its drivers (src/evp_v2b.F90:91-110, integration/ice_dyn_evp1d.F90:188-207).
 * Everything crosses as plain pointers, sizes and int status codes so that it
 * binds from Fortran (ISO_C_BINDING, see cice-evp1d_amd/fortran/evp1d_capi.F90),
 * C, C++ or Python ctypes.  No device pointer is ever handed to the caller: a
 * handle owns its device buffers and copies happen only in set/get calls.
 *
 * Conventions
 *   - neighbour indices are 1-based positions in 1..navel, as produced by
 *     findXinY (integration/ice_dyn_evp1d.F90:1000-1045);
 *   - skip flags are int32 (0 = active, nonzero = skip); Fortran logicals are
 *     converted at the boundary because their bit pattern is compiler specific;
 *   - all arithmetic is IEEE binary64; the default library is built with
 *     floating-point contraction off and reproduces the reference's serial
 *     operation order bit for bit.  libevp1d_fma.so exports the same ABI with
 *     contraction on (fused multiply-add);
 *   - every entry point returns EVP1D_OK (0) or a positive error code;
 *     evp1d_last_error() returns the message of the most recent failure.
 *   - one handle is driven from one host thread; work is issued on the
 *     handle's own HIP stream and every call that returns data synchronises it.
 */
#ifndef EVP1D_H
#define EVP1D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EVP1D_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum evp1d_status {
  EVP1D_OK = 0,
  EVP1D_EINVAL = 1,   /* bad argument (size, enum, null pointer)            */
  EVP1D_EINDEX = 2,   /* neighbour index outside 1..navel / inconsistent list */
  EVP1D_EHIP = 3,     /* HIP runtime error                                    */
  EVP1D_ENOMEM = 4,   /* device or host allocation failed                     */
  EVP1D_ESTATE = 5    /* call out of order (e.g. run before indices are set)  */
};

/* ---- scalar parameters --------------------------------------------------
 * Mirrors module ice_dyn_shared (src/ice.F90:80-107) plus the scalars that
 * readin_1d_v2 overrides from the input file (src/vars.F90:2116).  Field order
 * is part of the ABI (Fortran binds it as a bind(C) derived type).          */
typedef struct evp1d_params {
  int32_t ndte;        /* subcycles per model step (used by derived values)  */
  int32_t reserved;    /* keep 8-byte alignment of the doubles               */
  double capping;      /* replacement-pressure capping (1 = max form)        */
  double e_factor;     /* e_yieldcurve^2 / e_plasticpot^4                    */
  double epp2i;        /* 1 / e_plasticpot^2                                 */
  double deltaminEVP;  /* minimum Delta (times tarea)                        */
  double arlx;         /* 2 * elasticDamp * ndte                             */
  double arlx1i;       /* 1 / arlx                                           */
  double brlx;         /* ndte                                               */
  double denom1;       /* 1 / (1 + arlx1i)                                   */
  double revp;         /* 0 = classic EVP, 1 = revised EVP                   */
  double Ktens;        /* tensile strength factor                            */
  double u0;           /* residual velocity for basal stress (m/s)           */
  double cosw;         /* cos(ocean turning angle)                           */
  double sinw;         /* sin(ocean turning angle)                           */
  double rhow;         /* seawater density (kg/m^3)                          */
} evp1d_params;

/* set_evp_parameters (src/ice.F90:109-130): fills every member of *p from
 * ndte with the reference defaults (elasticDamp 0.36, e_yieldcurve =
 * e_plasticpot = 2, capping 1, deltaminEVP 1e-8, revp = Ktens = 0, u0 5e-5,
 * cosw 1, sinw 0, rhow 1026).  dt is accepted and unused, as in the reference. */
int evp1d_set_evp_parameters(int32_t ndte, double dt, evp1d_params *p);

/* ---- 1D fields ----------------------------------------------------------
 * Names follow src/vars.F90:185-199.  Fields marked [navel] have length navel
 * (active points 1..na then ghost slots na+1..navel); all others length na. */
enum evp1d_field {
  EVP1D_STRENGTH = 0,
  EVP1D_UVEL,          /* [navel] */
  EVP1D_VVEL,          /* [navel] */
  EVP1D_DXT, EVP1D_DYT, EVP1D_UAREAR,
  EVP1D_CDN_OCN,       /* Cw in stepu */
  EVP1D_AIU,           /* aiX in stepu */
  EVP1D_UOCN, EVP1D_VOCN, EVP1D_WATERX, EVP1D_WATERY,
  EVP1D_FORCEX, EVP1D_FORCEY, EVP1D_UMASSDTI, EVP1D_FM,
  EVP1D_STRINTX, EVP1D_STRINTY, EVP1D_TBU,
  EVP1D_STRESSP_1, EVP1D_STRESSP_2, EVP1D_STRESSP_3, EVP1D_STRESSP_4,
  EVP1D_STRESSM_1, EVP1D_STRESSM_2, EVP1D_STRESSM_3, EVP1D_STRESSM_4,
  EVP1D_STRESS12_1, EVP1D_STRESS12_2, EVP1D_STRESS12_3, EVP1D_STRESS12_4,
  EVP1D_HTE, EVP1D_HTN, EVP1D_HTEM1, EVP1D_HTNM1,
  EVP1D_UVEL_INIT, EVP1D_VVEL_INIT, EVP1D_CB,
  EVP1D_STR1, EVP1D_STR2, EVP1D_STR3, EVP1D_STR4,   /* [navel] */
  EVP1D_STR5, EVP1D_STR6, EVP1D_STR7, EVP1D_STR8,   /* [navel] */
  EVP1D_NFIELDS
};

/* neighbour lists (integration/ice_dyn_evp1d.F90:766-772, 792-797) */
enum evp1d_index { EVP1D_EE = 0, EVP1D_NE, EVP1D_SE, EVP1D_NW, EVP1D_SW, EVP1D_SSE };
/* skip flags (integration/ice_dyn_evp1d.F90:521-536) */
enum evp1d_skip { EVP1D_SKIPT = 0, EVP1D_SKIPU = 1 };
/* how evp1d_run keeps the subcycle loop on the device */
enum evp1d_mode {
  EVP1D_MODE_LAUNCH = 0,      /* one stress and one stepu launch per subcycle */
  EVP1D_MODE_GRAPH = 1,       /* the nsub x (stress, stepu, halo) sequence captured once as a hipGraph */
  EVP1D_MODE_PERSISTENT = 2,  /* one cooperative kernel for the whole loop: with the pair form selected
                                 (evp1d_set_fused_depth 2, planned) the pair form's tiles with one grid
                                 barrier per two subcycles, else stress and stepu over all points with a
                                 grid barrier between phases (hipLaunchCooperativeKernel) */
  EVP1D_MODE_AUTO = 3,        /* library default: FUSED when the lists tile and nsub >= 4, else GRAPH */
  EVP1D_MODE_FUSED = 4        /* one launch per subcycle: stress and stepu per tile of points, str1..8
                                 kept on chip (hipGraph); needs lists that describe a grid */
};

typedef struct evp1d_handle evp1d_handle;

/* Allocate a handle for na active points and navel-na ghost slots on the
 * current HIP device.  All state starts at zero (in particular str1..8 at the
 * ghost and skipped slots, which the stepu gathers read).  Replaces the
 * allocations of alloc_1d_v2 (src/vars.F90:2054-2100) and
 * evp1d_alloc_static_na/navel (integration/ice_dyn_evp1d.F90:378-461).     */
int evp1d_create(int32_t na, int32_t navel, const evp1d_params *p, evp1d_handle **out);
int evp1d_destroy(evp1d_handle *h);
int evp1d_set_params(evp1d_handle *h, const evp1d_params *p);
int evp1d_get_params(const evp1d_handle *h, evp1d_params *p);
int evp1d_get_sizes(const evp1d_handle *h, int32_t *na, int32_t *navel);

/* Host -> device copies.  n must equal the field's length.  Index lists are
 * validated (1 <= idx <= navel) before upload.                              */
int evp1d_set_field(evp1d_handle *h, int32_t field, const double *src, int32_t n);
int evp1d_get_field(evp1d_handle *h, int32_t field, double *dst, int32_t n);
int evp1d_set_index(evp1d_handle *h, int32_t which, const int32_t *idx, int32_t n);
int evp1d_set_skip(evp1d_handle *h, int32_t which, const int32_t *flags, int32_t n);

/* Ghost copy list: after every stepu, uvel/vvel of ghost slot dst[k] becomes
 * uvel/vvel of slot src[k] (src[k] == 0: fill with 0).  1-based; dst in
 * na+1..navel, src in 0..na.  This is the serial halo copy list of
 * src/vars.F90:100-143 (ice_haloUpdate) in 1D form.  n = 0 (the default)
 * reproduces the single-domain reference, whose ghost slots stay constant.  */
int evp1d_set_halo(evp1d_handle *h, int32_t n, const int32_t *dst, const int32_t *src);

/* One kernel each, over all na points (bench_v2b.F90 `stress` / `stepu` with
 * lb=1, ub=na; the ghost update `evp1d_halo` is a no-op with an empty list). */
int evp1d_stress(evp1d_handle *h);
int evp1d_stepu(evp1d_handle *h);
int evp1d_halo(evp1d_handle *h);

/* nsub subcycles of (stress; stepu; halo) without leaving the device
 * (src/evp_v2b.F90:91-110).  Asynchronous with respect to the host.          */
int evp1d_run(evp1d_handle *h, int32_t nsub);
int evp1d_set_mode(evp1d_handle *h, int32_t mode);
int evp1d_get_mode(const evp1d_handle *h, int32_t *mode);
/* nsub launches of stress only (which = 0) or stepu only (which = 1): the
 * lindividual timing loops of src/evp_v2b.F90:128-155.  which = 2: nsub full
 * subcycles of the fused form, one launch each, without a graph (for timing
 * the fused kernel; needs the fused form, see evp1d_set_tiling).             */
int evp1d_run_individual(evp1d_handle *h, int32_t which, int32_t nsub);
int evp1d_synchronize(evp1d_handle *h);

/* ---- ensembles: many domains in one handle ------------------------------
 * evp1d_create_batch concatenates nmem independent 1D lists (member k with
 * na[k] points and navel[k] slots) into one handle whose stress/stepu
 * launches cover every member at once.  Member k's active points occupy
 * caller positions a_off[k] .. a_off[k]+na[k]-1 (a_off = prefix sums of na),
 * its ghost slots follow all active points at positions ghost_off[k] ..
 * (ghost_off = sum(na) + prefix sums of navel-na), so the whole-handle calls
 * above see one list of sum(na) points and sum(navel) slots.  The *_member
 * calls take member-local arrays and 1-based member-local neighbour
 * positions (1..navel[k]) and apply the offsets; members never reference
 * each other's slots.  All members share the handle's scalar parameters.
 * This is the weak-scaling device of src/vars.F90:1794-2053 (replicate),
 * with the offsets applied consistently to active and ghost slots.         */
int evp1d_create_batch(int32_t nmem, const int32_t *na, const int32_t *navel, const evp1d_params *p,
                       evp1d_handle **out);
int evp1d_get_members(const evp1d_handle *h, int32_t *nmem);
int evp1d_member_info(const evp1d_handle *h, int32_t k, int32_t *na, int32_t *navel,
                      int32_t *active_offset, int32_t *ghost_offset);
int evp1d_set_field_member(evp1d_handle *h, int32_t k, int32_t field, const double *src, int32_t n);
int evp1d_get_field_member(evp1d_handle *h, int32_t k, int32_t field, double *dst, int32_t n);
int evp1d_set_index_member(evp1d_handle *h, int32_t k, int32_t which, const int32_t *idx, int32_t n);
int evp1d_set_skip_member(evp1d_handle *h, int32_t k, int32_t which, const int32_t *flags, int32_t n);

/* ---- 2D <-> 1D on the device ----------------------------------------------
 * Replace the host loops of convert_2d_1d_dyn and convert_1d_2d_dyn
 * (integration/ice_dyn_evp1d.F90:818-882, 884-947) for a single-domain
 * handle.  evp1d_set_grid gives the caller's 2D array size (nx x ny, Fortran
 * order, rim included) and indxtij (navel, 1-based linear index i + (j-1)*nx
 * of every 1D position, as evp1d_index_build returns it).
 * evp1d_put_2d copies nfield fields from `arrays` (nfield consecutive nx*ny
 * arrays) into the handle: position iw <= na takes G(indxti(iw), indxtj(iw));
 * for EVP1D_UVEL and EVP1D_VVEL the ghost positions become 0, as in the
 * reference (ice_dyn_evp1d.F90:880-881).  evp1d_get_2d writes nfield fields
 * back: G(indxti(iw), indxtj(iw)) = value for iw <= na; every other cell
 * takes the value that field's array had at its last evp1d_put_2d (so a
 * caller that has not changed the array in between sees it unchanged).  Both
 * return when the caller's arrays may be reused.                            */
int evp1d_set_grid(evp1d_handle *h, int32_t nx, int32_t ny, const int32_t *indxtij, int32_t navel);
/* Page-locked host memory for such arrays (transfers from it run at the full
 * host-link rate); free with evp1d_host_free.                               */
int evp1d_host_alloc(int64_t bytes, void **ptr);
int evp1d_host_free(void *ptr);
int evp1d_put_2d(evp1d_handle *h, int32_t nfield, const int32_t *fields, const double *arrays);
int evp1d_get_2d(evp1d_handle *h, int32_t nfield, const int32_t *fields, double *arrays);
/* evp1d_put_2d without its final wait (no counterpart in the reference): the
 * copies and gathers are queued on the handle's stream and the call returns
 * at once, so that the caller can fill its next array meanwhile.  arrays
 * must stay unchanged until a call that waits for the stream
 * (evp1d_synchronize, evp1d_get_2d, evp1d_get_field, evp1d_put_2d, ...),
 * and should be page-locked (evp1d_host_alloc) for the copies to overlap
 * the caller's work.                                                        */
int evp1d_put_2d_async(evp1d_handle *h, int32_t nfield, const int32_t *fields, const double *arrays);

/* ---- launch tuning (no counterpart in the reference) ---------------------
 * Block size (64, 128, 256, 512 or 1024 threads) and XCD-aware tile order of
 * the stress and stepu launches.  evp1d_autotune times every configuration
 * on the handle's own data (reps launches each; the state is saved and
 * restored) and keeps the fastest; it reports their average durations in
 * microseconds.  Results never depend on the tuning.                        */
int evp1d_autotune(evp1d_handle *h, int32_t reps, float *stress_us, float *stepu_us);
int evp1d_set_tuning(evp1d_handle *h, int32_t stress_block, int32_t stress_xcd, int32_t stepu_block,
                     int32_t stepu_xcd);
int evp1d_get_tuning(const evp1d_handle *h, int32_t *stress_block, int32_t *stress_xcd,
                     int32_t *stepu_block, int32_t *stepu_xcd);

/* ---- fused-subcycle tiling (no counterpart in the reference) -------------
 * The fused form runs stress and stepu of one subcycle for a tile of grid
 * cells in one workgroup: stress for the tile's own T points plus the T
 * points its U points read (evaluated again, never written), str1..8 kept in
 * LDS, then stepu.  Tiles are tile_w x tile_h boxes of grid coordinates that
 * the library recovers from the neighbour lists.  evp1d_set_tiling chooses
 * the tile ((tile_w+1)(tile_h+1) <= 2560), the workgroup size (64 to 512)
 * and the XCD-aware tile order (xcd = 0/1); a new tile re-lays the points out
 * at the next run.  Until it is called the library chooses: 31x19 cells on
 * 512 threads for the pair form (fused depth 2, the default), and for one
 * subcycle per launch by domain size, 64x16 cells on 512 threads if that
 * gives at least 1024 tiles, else 32x8 on 256 if at least 512, else 16x8 on
 * 128 (XCD-aware order).
 * evp1d_get_tiling reports the tiling and the number of tiles (0 when the
 * lists do not describe a grid: the fused form is then unavailable).
 * Results never depend on the tiling.                                       */
int evp1d_set_tiling(evp1d_handle *h, int32_t tile_w, int32_t tile_h, int32_t block, int32_t xcd);
int evp1d_get_tiling(evp1d_handle *h, int32_t *tile_w, int32_t *tile_h, int32_t *block, int32_t *xcd,
                     int32_t *ntiles);
/* Stress storage of fused runs.  inplace = 0 (default): stresses ping-pong
 * between two buffers like the velocities.  inplace = 1: each point's
 * stresses are updated in place by the workgroup that owns it; the points
 * that neighbouring tiles also evaluate go through a small double-buffered
 * copy during the run.  Which is faster depends on the domain size and the
 * cache hierarchy (in place moves the same bytes with a smaller footprint);
 * results are identical.                                                    */
int evp1d_set_fused_inplace(evp1d_handle *h, int32_t inplace);
int evp1d_get_fused_inplace(evp1d_handle *h, int32_t *inplace);
/* Subcycles per fused launch.  depth = 1: one.  depth = 2 (default; the
 * pair form): each workgroup runs two subcycles of its tile, the first over
 * a two-cell border ring that the neighbouring tiles also evaluate, so state
 * and constant fields cross HBM once per two subcycles; an odd nsub starts
 * with one one-subcycle launch.  The pair form uses two stress buffers, needs
 * (tile_w+1)(tile_h+1) <= 1024 on the lists and its tables within 160 KiB of
 * LDS, and picks its own workgroup size (one thread per level-2 stress
 * point); where it does not fit, fused runs use depth 1.  With a ghost
 * copy list (evp1d_set_halo) the copy also runs between the two subcycles
 * of a launch: a listed ghost read at the second subcycle takes its source's
 * value after the first.  evp1d_get_fused_depth reports
 * the setting and, in *active, whether runs use the pair form (0/1) and, in
 * *block, its workgroup size.  Results are identical.                       */
/* Order in which the fused forms visit their tiles, and so the slot order of
 * the points (no counterpart in the reference): 0 (default) rows of tiles,
 * 1 along a Hilbert curve over the tile grid (a space-filling-curve
 * ordering at tile granularity; with the XCD-aware order each XCD then takes
 * a compact block of tiles instead of a band).  Results are identical.     */
int evp1d_set_tile_order(evp1d_handle *h, int32_t order);
int evp1d_get_tile_order(evp1d_handle *h, int32_t *order);
int evp1d_set_fused_depth(evp1d_handle *h, int32_t depth);
int evp1d_get_fused_depth(evp1d_handle *h, int32_t *depth, int32_t *active, int32_t *block);
/* Cache hints of the pair form (no counterpart in the reference).  Its
 * launches can mark what each launch reads or writes once (stresses in and
 * out, the T points' constant fields, velocities and trajectory written) as
 * non-temporal loads and stores, so that the 256 MiB last-level cache keeps
 * what every launch reads again (the U points' constant fields and tables,
 * the velocities).  Same results bit for bit.  That pays when a launch's
 * bytes exceed that cache and the reread part fits it: 8 % less time per
 * subcycle at 1280x1520 (1.36M points) and at 1.68M; 9-13 % more time
 * below 0.7M points, 4-7 % more from 2.5M.  hints = -1 (default): chosen by
 * those sizes; 0: off; 1: on.  evp1d_get_cache_hints reports the setting and
 * whether the pair form's fused launches use them (pair form selected and
 * hints on; the persistent launch and the unfused forms never do).         */
int evp1d_set_cache_hints(evp1d_handle *h, int32_t hints);
int evp1d_get_cache_hints(evp1d_handle *h, int32_t *hints, int32_t *active);
/* Times one fused run of nsub subcycles (after an untimed one) for each tile
 * of 64x16/512, 32x16/512, 32x8/256, 16x8/128 (XCD-aware order) that fits
 * the lists, each with both stress storages, and the pair form on tiles of
 * 31x19, 27x23, 39x15, 33x17, 31x15, 62x19, 62x15, 15x19, 23x11, 15x15 and
 * 62x7 along the Hilbert curve and 31x19, 31x15, 15x15 in rows of tiles,
 * each without and with cache hints, on the handle's own data (the state is
 * saved and restored), keeps the fastest (as evp1d_set_tiling,
 * evp1d_set_fused_inplace, evp1d_set_fused_depth, evp1d_set_tile_order and
 * evp1d_set_cache_hints would) and reports its time per subcycle in
 * microseconds.                                                             */
int evp1d_autotune_fused(evp1d_handle *h, int32_t nsub, float *us);

/* Host-only tiling of a 1D list (no device use; for inspection and tests).
 * ee..sse: na each, 1-based slots as for evp1d_set_index; skipT, skipU: na
 * flags.  Outputs, each optional: perm[na] (slot -> 1-based point), counts[4]
 * (points that are stress only, stress and stepu, stepu only, neither),
 * *ntiles, *max_entries (largest LDS table), and, when the capacities
 * suffice, tiles[14 * ntiles] (a0,a1,b0,b1,c0,c1,h0,nh,k0,nk,sa,ia,sb,ib:
 * 0-based slot ranges of owned stress-only, owned stress-and-stepu and
 * stepu-only points, offset and length in list and in klist, then the owned
 * points other tiles read as halo: [a0, a0+sa) and [b0, b0+sb), numbered
 * from ia and ib in the run's table of such points), list[*nlist] (halo slots:
 * stress evaluated by the tile, written by their owner), klist[*nklist]
 * (constant slots: str1..8 read, never written) and uloc[4 * (counts[1] +
 * counts[2])] (LDS entries of own, nw, sse, sw for each stepu slot in slot
 * order; the LDS table is [owned stress-only | owned stress-and-stepu | halo |
 * constant]).  *ntiles = 0 when the lists do not describe a grid.  Returns
 * EVP1D_EINDEX if the plan fails its own consistency check.                  */
int evp1d_tile_plan(int32_t na, int32_t navel, const int32_t *ee, const int32_t *ne, const int32_t *se,
                    const int32_t *nw, const int32_t *sw, const int32_t *sse, const int32_t *skipT,
                    const int32_t *skipU, int32_t tile_w, int32_t tile_h, int32_t max_entries,
                    int32_t *perm, int32_t *counts, int32_t *ntiles, int32_t *max_used,
                    int32_t *tiles, int32_t tiles_cap, int32_t *list, int32_t list_cap, int32_t *nlist,
                    int32_t *klist, int32_t klist_cap, int32_t *nklist, int32_t *uloc);

/* ---- gradients (no counterpart in the reference) -------------------------
 * Loss L = 1/2 sum over stepu-active points (skipU = 0) of
 * (uvel - u_target)^2 + (vvel - v_target)^2 after nsub subcycles started from
 * the handle's current state.  evp1d_adjoint_run returns L and its gradient
 * with respect to strength, cdn_ocn (per point; the gradient with respect to
 * one scalar drag coefficient is the sum), forcex and forcey (caller order,
 * na each; any output may be NULL).  The reverse sweep uses hand-written
 * adjoint stress and stepu kernels and binomial checkpointing with nsnap
 * stored subcycle states (nsnap >= nsub stores them all: no recomputation,
 * 16 B per stepu point and subcycle; nsnap <= 0 chooses: all of them if they
 * fit in free device memory, else as many stored states as fit).  The handle
 * keeps the storage of either kind of call for later calls until it is
 * destroyed; a call with 0 < nsnap < nsub gives a whole trajectory back, and a
 * whole-trajectory call that has to allocate gives stored states back.
 * Afterwards the handle holds the forward state after nsub subcycles, exactly
 * as evp1d_run(h, nsub) leaves it.  evp1d_loss evaluates L at the current
 * state.                                                                    */
int evp1d_adjoint_run(evp1d_handle *h, int32_t nsub, int32_t nsnap, const double *u_target,
                      const double *v_target, double *loss, double *g_strength, double *g_cdn_ocn,
                      double *g_forcex, double *g_forcey);
int evp1d_loss(evp1d_handle *h, const double *u_target, const double *v_target, double *loss);
/* evp1d_loss without allocating (no counterpart in the reference): for
 * callers that cannot allocate device or host memory, such as tight
 * optimisation or serving loops.  The targets are staged in a fixed device
 * array of the library (64 MiB per GPU it runs on, held until the process
 * ends), so na may be at most 4,194,304 (EVP1D_EINVAL beyond).  The handle's
 * slot layout must be up to date (EVP1D_ESTATE after lists or skip flags
 * changed until evp1d_run, evp1d_loss or another call that uses the layout
 * has run).  L is summed in a fixed tree order on the device, so it can
 * differ from evp1d_loss's serial sum in the last bits.  All calls on one GPU
 * share the fixed array: calls that overlap on one GPU (other host threads or
 * handles) race.                                                            */
int evp1d_loss_noalloc(evp1d_handle *h, const double *u_target, const double *v_target, double *loss);
/* Checkpointing unit of evp1d_adjoint_run when 0 < nsnap < nsub (no
 * counterpart in the reference).  The run stores full states (stresses and
 * velocities) at unit starts by binomial checkpointing over units of seg
 * subcycles, and reverses each unit through a stored trajectory of its seg
 * subcycles' velocities (seg x 16 B per stepu point beside the nsnap stored
 * states).  seg = 1: units of one subcycle, reversed from str1..8 recomputed
 * from the stored state.  seg = 0 (default): ceil(nsub / nsnap), which stores
 * every unit start and recomputes each unit once, when that trajectory fits
 * in free device memory beside the stored states, else 1.  The gradient is
 * the same bit for bit for every seg and nsnap.                            */
int evp1d_set_adjoint_segment(evp1d_handle *h, int32_t seg);
int evp1d_get_adjoint_segment(evp1d_handle *h, int32_t *seg);
/* Device memory that gradient calls keep with the handle (no counterpart in
 * the reference): the stored trajectory, the stored states of checkpointed
 * calls and the adjoint work arrays.  evp1d_get_gradient_storage reports the
 * bytes held; evp1d_release_gradient_storage gives them back (the next
 * gradient call allocates again), for callers that need the memory for other
 * work without destroying the handle.                                       */
int evp1d_get_gradient_storage(evp1d_handle *h, int64_t *bytes);
int evp1d_release_gradient_storage(evp1d_handle *h);
/* Tangent-linear run: nsub subcycles from the current state propagating the
 * perturbation (d_strength, d_cdn_ocn, d_forcex, d_forcey) (caller order, na
 * each, NULL = zero) by forward-mode differentiation of the same point
 * functions; returns L and its directional derivative dL.  The handle ends
 * in the same state as evp1d_run(h, nsub).  With the adjoint gradient g,
 * dL = <g, d> (the dot-product test).                                       */
int evp1d_tangent_run(evp1d_handle *h, int32_t nsub, const double *d_strength, const double *d_cdn_ocn,
                      const double *d_forcex, const double *d_forcey, const double *u_target,
                      const double *v_target, double *loss, double *dloss);

/* Device timing on the handle's stream (HIP events): start, then stop
 * returns the elapsed milliseconds between the two after synchronising.     */
int evp1d_timer_start(evp1d_handle *h);
int evp1d_timer_stop(evp1d_handle *h, float *ms);

/* Build the 1D point list and neighbour positions from a 2D T mask, as
 * calc_nActiveTU / calc_2d_indices_init / calc_navel / convert_2d_1d_init
 * do (integration/ice_dyn_evp1d.F90:494-519, 559-589, 743-816, 1047-1087).
 * nx, ny are the full block sizes including the ghost rim (cnx + 2*nghost);
 * tmask is int32 (nonzero = ocean), Fortran order (i fastest).  The rim cells
 * with i or j <= nghost must be dry (the reference silently leaves zero
 * indices there).  Outputs: indxti/indxtj (na, 1-based 2D indices),
 * indxtij (navel, 1-based linear index i + (j-1)*nx of every 1D slot; ghost
 * slots are the sorted set difference), and ee..sse (na each, 1-based).     */
int evp1d_index_count(int32_t nx, int32_t ny, int32_t nghost, const int32_t *tmask,
                      int32_t *na, int32_t *navel);
int evp1d_index_build(int32_t nx, int32_t ny, int32_t nghost, const int32_t *tmask,
                      int32_t na, int32_t navel, int32_t *indxti, int32_t *indxtj,
                      int32_t *indxtij, int32_t *ee, int32_t *ne, int32_t *se,
                      int32_t *nw, int32_t *sw, int32_t *sse);

/* ---- the v0 2D block path --------------------------------------------------
 * The reference's v0 benchmark keeps CICE's own layout: every field is an
 * (nx_block, ny_block, max_blocks) array in Fortran order, and each block has
 * a list of T cells for stress and of U cells for stepu (src/evp_v0.F90:93-136,
 * src/bench_v0.F90:62-82 stress, 469-486 stepu).  An evp2d handle holds those
 * arrays on the device and runs the same subcycles; results are bit-identical
 * to the reference v0 (and, on the same grid, to the 1D path).
 *
 * Fields, in the order of input_double_2d.bin (src/vars.F90:299-305) and then
 * the arrays the v0 driver keeps besides; str1..8 are stress's scratch (the
 * reference clears them in every stress call, bench_v0.F90:135), so they can
 * be read but not set.                                                      */
enum evp2d_field {
  EVP2D_STRENGTH = 0, EVP2D_UVEL, EVP2D_VVEL, EVP2D_DXT, EVP2D_DYT, EVP2D_DXHY, EVP2D_DYHX,
  EVP2D_CXP, EVP2D_CYP, EVP2D_CXM, EVP2D_CYM, EVP2D_DMINTAREA, EVP2D_UAREAR, EVP2D_CDN_OCN,
  EVP2D_AIX, EVP2D_UOCN, EVP2D_VOCN, EVP2D_WATERX, EVP2D_WATERY, EVP2D_FORCEX, EVP2D_FORCEY,
  EVP2D_UMASSDTI, EVP2D_FM, EVP2D_STRINTX, EVP2D_STRINTY, EVP2D_TBU,
  EVP2D_STRESSP_1, EVP2D_STRESSP_2, EVP2D_STRESSP_3, EVP2D_STRESSP_4,
  EVP2D_STRESSM_1, EVP2D_STRESSM_2, EVP2D_STRESSM_3, EVP2D_STRESSM_4,
  EVP2D_STRESS12_1, EVP2D_STRESS12_2, EVP2D_STRESS12_3, EVP2D_STRESS12_4,
  EVP2D_UVEL_INIT, EVP2D_VVEL_INIT, EVP2D_TAUBX, EVP2D_TAUBY,
  EVP2D_STR1, EVP2D_STR2, EVP2D_STR3, EVP2D_STR4, EVP2D_STR5, EVP2D_STR6, EVP2D_STR7, EVP2D_STR8,
  EVP2D_NFIELDS
};

typedef struct evp2d_handle evp2d_handle;

/* alloc_2d_v0 (src/vars.F90:222-278): arrays of nx_block x ny_block x
 * max_blocks on the device, all zero; params as for evp1d_create.           */
int evp2d_create(int32_t nx_block, int32_t ny_block, int32_t max_blocks, const evp1d_params *p,
                 evp2d_handle **out);
int evp2d_destroy(evp2d_handle *h);
int evp2d_set_params(evp2d_handle *h, const evp1d_params *p);
/* The icellt/indxti/indxtj and icellu/indxui/indxuj arguments of stress and
 * stepu for block iblk (1-based; bench_v0.F90:62-76, 469-482), 1-based cell
 * coordinates.  T cells need 2 <= i <= nx_block, 2 <= j <= ny_block (stress
 * reads i-1, j-1), U cells i < nx_block, j < ny_block (stepu reads i+1,
 * j+1); a cell listed twice is rejected (EVP1D_EINDEX).                      */
int evp2d_set_cells(evp2d_handle *h, int32_t iblk, int32_t icellt, const int32_t *indxti,
                    const int32_t *indxtj, int32_t icellu, const int32_t *indxui, const int32_t *indxuj);
/* readin_halo's tables (src/vars.F90:80-99): per local block the interior
 * bounds ilo, jlo, ihi, jhi, and ncopies copies src[3k..3k+2] = (i, j, block)
 * -> dst[3k..3k+2]; source block 0 means fill with 0.  Applied to uvel and
 * vvel after every subcycle when max_blocks > 1 (ice_haloUpdate, 100-143):
 * the halo rows and columns of each block are set to 0, then the copies run.
 * A copy whose source cell is written by another copy (whose result would
 * depend on the list order) is rejected (EVP1D_EINVAL).                     */
int evp2d_set_halo(evp2d_handle *h, int32_t nblocks, const int32_t *ilo, const int32_t *jlo,
                   const int32_t *ihi, const int32_t *jhi, int32_t ncopies, const int32_t *src,
                   const int32_t *dst);
/* Whole arrays, n = nx_block * ny_block * max_blocks values, Fortran order. */
int evp2d_set_field(evp2d_handle *h, int32_t field, const double *a, int64_t n);
int evp2d_get_field(evp2d_handle *h, int32_t field, double *a, int64_t n);
/* nsub subcycles: stress and stepu of every block, then the halo update when
 * max_blocks > 1 and tables are set (src/evp_v0.F90:93-136); a captured
 * hipGraph, asynchronous.                                                   */
int evp2d_run(evp2d_handle *h, int32_t nsub);
/* How evp2d_run runs a subcycle: EVP2D_MODE_FUSED (default) runs stress and
 * stepu of one 32 x 8 tile of cells per workgroup with str1..8 in LDS (the
 * T cells of the column east and the row north of the tile evaluated again),
 * velocities and stresses alternating between two buffers; EVP2D_MODE_PLAIN
 * runs the two kernels one after the other, as the reference does.  Results
 * are identical.                                                            */
enum evp2d_mode { EVP2D_MODE_PLAIN = 0, EVP2D_MODE_FUSED = 1 };
int evp2d_set_mode(evp2d_handle *h, int32_t mode);
int evp2d_get_mode(evp2d_handle *h, int32_t *mode);
/* Cache hints of the fused form (no counterpart in the reference), as
 * evp1d_set_cache_hints: the tile kernel reads and writes its packed
 * stresses (and the last subcycle's str1..8) as non-temporal loads and
 * stores, so that the last-level cache keeps the packed constants, tables
 * and velocities between subcycles.  Same results bit for bit.  hints = -1
 * (default): chosen by the domain size; 0: off; 1: on.
 * evp2d_get_cache_hints reports the setting and whether fused runs use
 * them.                                                                     */
int evp2d_set_cache_hints(evp2d_handle *h, int32_t hints);
int evp2d_get_cache_hints(evp2d_handle *h, int32_t *hints, int32_t *active);
int evp2d_stress(evp2d_handle *h);
int evp2d_stepu(evp2d_handle *h);
int evp2d_halo(evp2d_handle *h);
/* the lindividual loops of evp_v0.F90:157-204: which = 0 stress, 1 stepu;
 * which = 2 runs nsub fused subcycles launched one by one (timing)          */
int evp2d_run_individual(evp2d_handle *h, int32_t which, int32_t nsub);
int evp2d_synchronize(evp2d_handle *h);
int evp2d_timer_start(evp2d_handle *h);
int evp2d_timer_stop(evp2d_handle *h, float *ms);
/* total T and U cells over all blocks                                       */
int evp2d_get_cells(evp2d_handle *h, int32_t *nt, int32_t *nu);

/* Copy the message of the last failure on this thread into buf (always NUL
 * terminated); returns the code of that failure (0 if none).                */
int evp1d_last_error(char *buf, int32_t n);
/* ABI version and build flavour: fma = 1 for libevp1d_fma.so.               */
int evp1d_build_info(int32_t *abi_version, int32_t *fma);

#ifdef __cplusplus
}
#endif
#endif /* EVP1D_H */
