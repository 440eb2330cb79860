!===============================================================================
! evp1d_capi - ISO_C_BINDING interface to libevp1d.so (include/evp1d.h).
!
! One interface block per C entry point, same names.  The reference's only
! Fortran<->C precedent is module ssc_fortran (src/evp_v2b.F90:9-21 binding
! src/api_fortran_sde.c:21-29); this module uses the same mechanism with
! array and scalar arguments.  evp1d_check() turns a nonzero status into the
! reference's error convention: print the message and stop.
!===============================================================================
module evp1d_capi
  use, intrinsic :: iso_c_binding
  implicit none
  private

  ! evp1d_params, field for field (include/evp1d.h)
  type, bind(C), public :: evp1d_params_t
    integer(c_int32_t) :: ndte = 0, reserved = 0
    real(c_double) :: capping = 0, e_factor = 0, epp2i = 0, deltaminEVP = 0
    real(c_double) :: arlx = 0, arlx1i = 0, brlx = 0, denom1 = 0
    real(c_double) :: revp = 0, Ktens = 0, u0 = 0, cosw = 0, sinw = 0, rhow = 0
  end type evp1d_params_t

  ! status codes
  integer(c_int), parameter, public :: EVP1D_OK = 0, EVP1D_EINVAL = 1, EVP1D_EINDEX = 2, &
                                       EVP1D_EHIP = 3, EVP1D_ENOMEM = 4, EVP1D_ESTATE = 5
  ! fields (enum evp1d_field)
  integer(c_int32_t), parameter, public :: &
    F_STRENGTH = 0, F_UVEL = 1, F_VVEL = 2, F_DXT = 3, F_DYT = 4, F_UAREAR = 5, &
    F_CDN_OCN = 6, F_AIU = 7, F_UOCN = 8, F_VOCN = 9, F_WATERX = 10, F_WATERY = 11, &
    F_FORCEX = 12, F_FORCEY = 13, F_UMASSDTI = 14, F_FM = 15, F_STRINTX = 16, &
    F_STRINTY = 17, F_TBU = 18, &
    F_STRESSP_1 = 19, F_STRESSP_2 = 20, F_STRESSP_3 = 21, F_STRESSP_4 = 22, &
    F_STRESSM_1 = 23, F_STRESSM_2 = 24, F_STRESSM_3 = 25, F_STRESSM_4 = 26, &
    F_STRESS12_1 = 27, F_STRESS12_2 = 28, F_STRESS12_3 = 29, F_STRESS12_4 = 30, &
    F_HTE = 31, F_HTN = 32, F_HTEM1 = 33, F_HTNM1 = 34, F_UVEL_INIT = 35, &
    F_VVEL_INIT = 36, F_CB = 37, &
    F_STR1 = 38, F_STR2 = 39, F_STR3 = 40, F_STR4 = 41, &
    F_STR5 = 42, F_STR6 = 43, F_STR7 = 44, F_STR8 = 45
  ! neighbour lists, skip flags, modes
  integer(c_int32_t), parameter, public :: I_EE = 0, I_NE = 1, I_SE = 2, I_NW = 3, &
                                           I_SW = 4, I_SSE = 5
  integer(c_int32_t), parameter, public :: S_SKIPT = 0, S_SKIPU = 1
  integer(c_int32_t), parameter, public :: MODE_LAUNCH = 0, MODE_GRAPH = 1, &
                                           MODE_PERSISTENT = 2, MODE_AUTO = 3, MODE_FUSED = 4
  ! fields of the v0 2D block path (enum evp2d_field): input_double_2d.bin's
  ! order (src/vars.F90:299-305), then the driver's own arrays
  integer(c_int32_t), parameter, public :: &
    V_STRENGTH = 0, V_UVEL = 1, V_VVEL = 2, V_DXT = 3, V_DYT = 4, V_DXHY = 5, V_DYHX = 6, &
    V_CXP = 7, V_CYP = 8, V_CXM = 9, V_CYM = 10, V_DMINTAREA = 11, V_UAREAR = 12, &
    V_CDN_OCN = 13, V_AIX = 14, V_UOCN = 15, V_VOCN = 16, V_WATERX = 17, V_WATERY = 18, &
    V_FORCEX = 19, V_FORCEY = 20, V_UMASSDTI = 21, V_FM = 22, V_STRINTX = 23, &
    V_STRINTY = 24, V_TBU = 25, V_STRESSP_1 = 26, V_STRESS12_4 = 37, V_UVEL_INIT = 38, &
    V_VVEL_INIT = 39, V_TAUBX = 40, V_TAUBY = 41, V_STR1 = 42, V_STR8 = 49
  integer(c_int32_t), parameter, public :: MODE2D_PLAIN = 0, MODE2D_FUSED = 1

  public :: evp1d_set_evp_parameters, evp1d_create, evp1d_destroy, evp1d_set_params,   &
            evp1d_get_params, evp1d_get_sizes, evp1d_set_field, evp1d_get_field,        &
            evp1d_set_index, evp1d_set_skip, evp1d_set_halo, evp1d_stress, evp1d_stepu, &
            evp1d_halo, evp1d_run, evp1d_set_mode, evp1d_get_mode,                      &
            evp1d_run_individual, evp1d_synchronize, evp1d_timer_start,                 &
            evp1d_timer_stop, evp1d_index_count, evp1d_index_build, evp1d_last_error,   &
            evp1d_build_info, evp1d_check, evp1d_create_batch, evp1d_get_members,       &
            evp1d_member_info, evp1d_set_field_member, evp1d_get_field_member,          &
            evp1d_set_index_member, evp1d_set_skip_member, evp1d_autotune,              &
            evp1d_set_tuning, evp1d_get_tuning, evp1d_set_tiling, evp1d_get_tiling,       &
            evp1d_set_grid, evp1d_put_2d, evp1d_put_2d_async, evp1d_get_2d, evp1d_host_alloc, &
            evp1d_host_free, &
            evp1d_set_fused_inplace, evp1d_get_fused_inplace, evp1d_autotune_fused,    &
            evp1d_set_fused_depth, evp1d_get_fused_depth,                                &
            evp1d_set_cache_hints, evp1d_get_cache_hints,                                &
            evp1d_adjoint_run, evp1d_loss, evp1d_tangent_run, evp1d_set_adjoint_segment, &
            evp1d_get_adjoint_segment, evp1d_set_tile_order, evp1d_get_tile_order,       &
            evp1d_get_gradient_storage, evp1d_release_gradient_storage,                  &
            evp2d_create, evp2d_destroy, evp2d_set_params, evp2d_set_cells, evp2d_set_halo, &
            evp2d_set_field, evp2d_get_field, evp2d_run, evp2d_stress, evp2d_stepu,      &
            evp2d_halo, evp2d_run_individual, evp2d_synchronize, evp2d_timer_start,    &
            evp2d_timer_stop, evp2d_get_cells, evp2d_set_mode, evp2d_get_mode,           &
            evp2d_set_cache_hints, evp2d_get_cache_hints

  interface
    integer(c_int) function evp1d_set_evp_parameters(ndte, dt, p) bind(C, name='evp1d_set_evp_parameters')
      import :: c_int, c_int32_t, c_double, evp1d_params_t
      integer(c_int32_t), value :: ndte
      real(c_double), value :: dt
      type(evp1d_params_t), intent(out) :: p
    end function
    integer(c_int) function evp1d_create(na, navel, p, h) bind(C, name='evp1d_create')
      import :: c_int, c_int32_t, c_ptr, evp1d_params_t
      integer(c_int32_t), value :: na, navel
      type(evp1d_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function evp1d_destroy(h) bind(C, name='evp1d_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp1d_set_params(h, p) bind(C, name='evp1d_set_params')
      import :: c_int, c_ptr, evp1d_params_t
      type(c_ptr), value :: h
      type(evp1d_params_t), intent(in) :: p
    end function
    integer(c_int) function evp1d_get_params(h, p) bind(C, name='evp1d_get_params')
      import :: c_int, c_ptr, evp1d_params_t
      type(c_ptr), value :: h
      type(evp1d_params_t), intent(out) :: p
    end function
    integer(c_int) function evp1d_get_sizes(h, na, navel) bind(C, name='evp1d_get_sizes')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: na, navel
    end function
    integer(c_int) function evp1d_set_field(h, field, src, n) bind(C, name='evp1d_set_field')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: field, n
      real(c_double), intent(in) :: src(*)
    end function
    integer(c_int) function evp1d_get_field(h, field, dst, n) bind(C, name='evp1d_get_field')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: field, n
      real(c_double), intent(out) :: dst(*)
    end function
    integer(c_int) function evp1d_set_index(h, which, idx, n) bind(C, name='evp1d_set_index')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: which, n
      integer(c_int32_t), intent(in) :: idx(*)
    end function
    integer(c_int) function evp1d_set_skip(h, which, flags, n) bind(C, name='evp1d_set_skip')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: which, n
      integer(c_int32_t), intent(in) :: flags(*)
    end function
    integer(c_int) function evp1d_set_halo(h, n, dst, src) bind(C, name='evp1d_set_halo')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: dst(*), src(*)
    end function
    integer(c_int) function evp1d_stress(h) bind(C, name='evp1d_stress')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp1d_stepu(h) bind(C, name='evp1d_stepu')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp1d_halo(h) bind(C, name='evp1d_halo')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp1d_run(h, nsub) bind(C, name='evp1d_run')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsub
    end function
    integer(c_int) function evp1d_set_mode(h, mode) bind(C, name='evp1d_set_mode')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: mode
    end function
    integer(c_int) function evp1d_get_mode(h, mode) bind(C, name='evp1d_get_mode')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: mode
    end function
    integer(c_int) function evp1d_run_individual(h, which, nsub) bind(C, name='evp1d_run_individual')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: which, nsub
    end function
    integer(c_int) function evp1d_synchronize(h) bind(C, name='evp1d_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp1d_timer_start(h) bind(C, name='evp1d_timer_start')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp1d_timer_stop(h, ms) bind(C, name='evp1d_timer_stop')
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: h
      real(c_float), intent(out) :: ms
    end function
    integer(c_int) function evp1d_index_count(nx, ny, nghost, tmask, na, navel) bind(C, name='evp1d_index_count')
      import :: c_int, c_int32_t
      integer(c_int32_t), value :: nx, ny, nghost
      integer(c_int32_t), intent(in) :: tmask(*)
      integer(c_int32_t), intent(out) :: na, navel
    end function
    integer(c_int) function evp1d_index_build(nx, ny, nghost, tmask, na, navel, indxti, indxtj, indxtij, &
                                              ee, ne, se, nw, sw, sse) bind(C, name='evp1d_index_build')
      import :: c_int, c_int32_t
      integer(c_int32_t), value :: nx, ny, nghost, na, navel
      integer(c_int32_t), intent(in) :: tmask(*)
      integer(c_int32_t), intent(out) :: indxti(*), indxtj(*), indxtij(*), ee(*), ne(*), se(*), &
                                         nw(*), sw(*), sse(*)
    end function
    integer(c_int) function evp1d_set_grid(h, nx, ny, indxtij, navel) bind(C, name='evp1d_set_grid')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nx, ny, navel
      integer(c_int32_t), intent(in) :: indxtij(*)
    end function
    integer(c_int) function evp1d_put_2d(h, nfield, fields, arrays) bind(C, name='evp1d_put_2d')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nfield
      integer(c_int32_t), intent(in) :: fields(*)
      real(c_double), intent(in) :: arrays(*)
    end function
    ! arrays must stay unchanged until a call that waits (see include/evp1d.h):
    ! pass page-locked memory that outlives the queued copies
    integer(c_int) function evp1d_put_2d_async(h, nfield, fields, arrays) bind(C, name='evp1d_put_2d_async')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nfield
      integer(c_int32_t), intent(in) :: fields(*)
      real(c_double), intent(in) :: arrays(*)
    end function
    integer(c_int) function evp1d_get_2d(h, nfield, fields, arrays) bind(C, name='evp1d_get_2d')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nfield
      integer(c_int32_t), intent(in) :: fields(*)
      real(c_double), intent(inout) :: arrays(*)
    end function
    integer(c_int) function evp1d_host_alloc(bytes, ptr) bind(C, name='evp1d_host_alloc')
      import :: c_int, c_int64_t, c_ptr
      integer(c_int64_t), value :: bytes
      type(c_ptr), intent(out) :: ptr
    end function
    integer(c_int) function evp1d_host_free(ptr) bind(C, name='evp1d_host_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
    end function
    integer(c_int) function evp1d_last_error(buf, n) bind(C, name='evp1d_last_error')
      import :: c_int, c_int32_t, c_char
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int32_t), value :: n
    end function
    integer(c_int) function evp1d_build_info(abi, fma) bind(C, name='evp1d_build_info')
      import :: c_int, c_int32_t
      integer(c_int32_t), intent(out) :: abi, fma
    end function
    integer(c_int) function evp1d_create_batch(nmem, na, navel, p, h) bind(C, name='evp1d_create_batch')
      import :: c_int, c_int32_t, c_ptr, evp1d_params_t
      integer(c_int32_t), value :: nmem
      integer(c_int32_t), intent(in) :: na(*), navel(*)
      type(evp1d_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function evp1d_get_members(h, nmem) bind(C, name='evp1d_get_members')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: nmem
    end function
    integer(c_int) function evp1d_member_info(h, k, na, navel, aoff, goff) bind(C, name='evp1d_member_info')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: k
      integer(c_int32_t), intent(out) :: na, navel, aoff, goff
    end function
    integer(c_int) function evp1d_set_field_member(h, k, field, src, n) bind(C, name='evp1d_set_field_member')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: k, field, n
      real(c_double), intent(in) :: src(*)
    end function
    integer(c_int) function evp1d_get_field_member(h, k, field, dst, n) bind(C, name='evp1d_get_field_member')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: k, field, n
      real(c_double), intent(out) :: dst(*)
    end function
    integer(c_int) function evp1d_set_index_member(h, k, which, idx, n) bind(C, name='evp1d_set_index_member')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: k, which, n
      integer(c_int32_t), intent(in) :: idx(*)
    end function
    integer(c_int) function evp1d_set_skip_member(h, k, which, flags, n) bind(C, name='evp1d_set_skip_member')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: k, which, n
      integer(c_int32_t), intent(in) :: flags(*)
    end function
    integer(c_int) function evp1d_autotune(h, reps, stress_us, stepu_us) bind(C, name='evp1d_autotune')
      import :: c_int, c_int32_t, c_float, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: reps
      real(c_float), intent(out) :: stress_us, stepu_us
    end function
    integer(c_int) function evp1d_set_tuning(h, sb, sx, ub, ux) bind(C, name='evp1d_set_tuning')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: sb, sx, ub, ux
    end function
    integer(c_int) function evp1d_get_tuning(h, sb, sx, ub, ux) bind(C, name='evp1d_get_tuning')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: sb, sx, ub, ux
    end function
    integer(c_int) function evp1d_set_tiling(h, tw, th, block, xcd) bind(C, name='evp1d_set_tiling')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: tw, th, block, xcd
    end function
    integer(c_int) function evp1d_get_tiling(h, tw, th, block, xcd, ntiles) bind(C, name='evp1d_get_tiling')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: tw, th, block, xcd, ntiles
    end function
    integer(c_int) function evp1d_set_fused_inplace(h, inplace) bind(C, name='evp1d_set_fused_inplace')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: inplace
    end function
    integer(c_int) function evp1d_get_fused_inplace(h, inplace) bind(C, name='evp1d_get_fused_inplace')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: inplace
    end function
    integer(c_int) function evp1d_set_fused_depth(h, depth) bind(C, name='evp1d_set_fused_depth')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: depth
    end function
    integer(c_int) function evp1d_get_fused_depth(h, depth, active, block) bind(C, name='evp1d_get_fused_depth')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: depth, active, block
    end function
    integer(c_int) function evp1d_set_cache_hints(h, hints) bind(C, name='evp1d_set_cache_hints')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: hints
    end function
    integer(c_int) function evp1d_get_cache_hints(h, hints, active) bind(C, name='evp1d_get_cache_hints')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: hints, active
    end function
    ! ---- the v0 2D block path -------------------------------------------
    integer(c_int) function evp2d_create(nxb, nyb, nblocks, p, h) bind(C, name='evp2d_create')
      import :: c_int, c_int32_t, c_ptr, evp1d_params_t
      integer(c_int32_t), value :: nxb, nyb, nblocks
      type(evp1d_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function evp2d_destroy(h) bind(C, name='evp2d_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp2d_set_params(h, p) bind(C, name='evp2d_set_params')
      import :: c_int, c_ptr, evp1d_params_t
      type(c_ptr), value :: h
      type(evp1d_params_t), intent(in) :: p
    end function
    integer(c_int) function evp2d_set_cells(h, iblk, icellt, indxti, indxtj, icellu, indxui, indxuj) &
        bind(C, name='evp2d_set_cells')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: iblk, icellt, icellu
      integer(c_int32_t), intent(in) :: indxti(*), indxtj(*), indxui(*), indxuj(*)
    end function
    integer(c_int) function evp2d_set_halo(h, nblocks, ilo, jlo, ihi, jhi, ncopies, src, dst) &
        bind(C, name='evp2d_set_halo')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nblocks, ncopies
      integer(c_int32_t), intent(in) :: ilo(*), jlo(*), ihi(*), jhi(*), src(*), dst(*)
    end function
    integer(c_int) function evp2d_set_field(h, field, a, n) bind(C, name='evp2d_set_field')
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: field
      real(c_double), intent(in) :: a(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function evp2d_get_field(h, field, a, n) bind(C, name='evp2d_get_field')
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: field
      real(c_double), intent(out) :: a(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function evp2d_run(h, nsub) bind(C, name='evp2d_run')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsub
    end function
    integer(c_int) function evp2d_stress(h) bind(C, name='evp2d_stress')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp2d_stepu(h) bind(C, name='evp2d_stepu')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp2d_halo(h) bind(C, name='evp2d_halo')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp2d_run_individual(h, which, nsub) bind(C, name='evp2d_run_individual')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: which, nsub
    end function
    integer(c_int) function evp2d_synchronize(h) bind(C, name='evp2d_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp2d_timer_start(h) bind(C, name='evp2d_timer_start')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp2d_timer_stop(h, ms) bind(C, name='evp2d_timer_stop')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), intent(out) :: ms
    end function
    integer(c_int) function evp2d_set_mode(h, mode) bind(C, name='evp2d_set_mode')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: mode
    end function
    integer(c_int) function evp2d_get_mode(h, mode) bind(C, name='evp2d_get_mode')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: mode
    end function
    integer(c_int) function evp2d_set_cache_hints(h, hints) bind(C, name='evp2d_set_cache_hints')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: hints
    end function
    integer(c_int) function evp2d_get_cache_hints(h, hints, active) bind(C, name='evp2d_get_cache_hints')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: hints, active
    end function
    integer(c_int) function evp2d_get_cells(h, nt, nu) bind(C, name='evp2d_get_cells')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: nt, nu
    end function
    ! gradients (include/evp1d.h): arrays in caller order, na each; every
    ! output array must be present (the C side also accepts NULL)
    integer(c_int) function evp1d_adjoint_run(h, nsub, nsnap, u_target, v_target, loss, g_strength, &
                                              g_cdn_ocn, g_forcex, g_forcey) bind(C, name='evp1d_adjoint_run')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsub, nsnap
      real(c_double), intent(in) :: u_target(*), v_target(*)
      real(c_double), intent(out) :: loss
      real(c_double), intent(out) :: g_strength(*), g_cdn_ocn(*), g_forcex(*), g_forcey(*)
    end function
    integer(c_int) function evp1d_loss(h, u_target, v_target, loss) bind(C, name='evp1d_loss')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u_target(*), v_target(*)
      real(c_double), intent(out) :: loss
    end function
    integer(c_int) function evp1d_tangent_run(h, nsub, d_strength, d_cdn_ocn, d_forcex, d_forcey, u_target, &
                                              v_target, loss, dloss) bind(C, name='evp1d_tangent_run')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsub
      real(c_double), intent(in) :: d_strength(*), d_cdn_ocn(*), d_forcex(*), d_forcey(*)
      real(c_double), intent(in) :: u_target(*), v_target(*)
      real(c_double), intent(out) :: loss, dloss
    end function
    integer(c_int) function evp1d_set_adjoint_segment(h, seg) bind(C, name='evp1d_set_adjoint_segment')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: seg
    end function
    integer(c_int) function evp1d_get_adjoint_segment(h, seg) bind(C, name='evp1d_get_adjoint_segment')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: seg
    end function
    integer(c_int) function evp1d_get_gradient_storage(h, bytes) bind(C, name='evp1d_get_gradient_storage')
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: bytes
    end function
    integer(c_int) function evp1d_release_gradient_storage(h) bind(C, name='evp1d_release_gradient_storage')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function evp1d_set_tile_order(h, order) bind(C, name='evp1d_set_tile_order')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: order
    end function
    integer(c_int) function evp1d_get_tile_order(h, order) bind(C, name='evp1d_get_tile_order')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: order
    end function
    integer(c_int) function evp1d_autotune_fused(h, nsub, us) bind(C, name='evp1d_autotune_fused')
      import :: c_int, c_int32_t, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsub
      real(c_float), intent(out) :: us
    end function
  end interface

contains

  ! stop with the library's message when rc /= 0 (the reference stops on
  ! allocation and consistency errors, e.g. src/vars.F90:2063)
  subroutine evp1d_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, code
    if (rc == EVP1D_OK) return
    code = evp1d_last_error(buf, 512_c_int32_t)
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    enddo
    write(*,'(a,a,a,i0,a,a)') 'evp1d error in ', trim(where), ' (status ', rc, '): ', trim(msg)
    stop 1
  end subroutine evp1d_check

end module evp1d_capi
