!===============================================================================
! ice_dyn_evp1d - CICE-facing evp1d driver API on the MI355X engine.
!
! Same public names and argument order as the reference's integration module
! (integration/ice_dyn_evp1d.F90:16, 65-77, 115-138, 258-266, 326):
!   dyn_evp1d_init      gather the static 2D fields, build the 1D point list
!                       and neighbour positions, upload the static metrics
!   dyn_evp1d_run       gather the dynamic 2D fields, set the skip flags from
!                       the ice masks, pack 2D -> 1D, run ndte subcycles on the
!                       GPU, unpack 1D -> 2D and scatter back to the caller
!   dyn_evp2d_dump      write the 29 dynamic 2D fields to a stream file
!   dyn_evp1d_finalize  release the GPU state and close the log
! Differences from the reference, each deliberate:
!   * the arrays the subcycles change (12 stresses, strintx/y, Cb, uvel, vvel)
!     are intent(inout) and receive the results; the reference declares them
!     intent(in) and unpacks into locals, so its results never reach the
!     caller (integration/ice_dyn_evp1d.F90:129-138, 217-225);
!   * str1..8 start at zero (the reference leaves them uninitialised,
!     ice_dyn_evp1d.F90:455-459) and uvel_init/vvel_init are the packed
!     uvel/vvel of the run (the reference never sets them);
!   * wet cells in the west/south ghost rim stop with a message instead of
!     leaving zero indices (ice_dyn_evp1d.F90:503-517 vs 579-580).
! The subcycle loop, the 1D list construction and the neighbour search run in
! libevp1d.so (include/evp1d.h) through module evp1d_capi.  With ldevpack
! (namelist evp1d_nml, default .true.) the 2D <-> 1D copies of the dynamic
! fields (convert_2d_1d_dyn / convert_1d_2d_dyn, ice_dyn_evp1d.F90:818-947)
! also run on the device: whole 2D arrays cross PCIe and the library gathers
! and scatters them; .false. keeps the host loops.  With lautotune (default
! .false.) the first run first times the fused form's tiles, stress
! storages, depth and cache hints on the first step's own fields
! (evp1d_autotune_fused, state restored) and keeps the fastest, written to
! evp1d.log with linfo; results do not depend on it.
!===============================================================================
module ice_dyn_evp1d
  use, intrinsic :: iso_c_binding
  use ice_kinds_mod
  use ice_constants
  use evp1d_capi
  implicit none
  private

  public :: dyn_evp1d_init, dyn_evp1d_run, dyn_evp1d_finalize, dyn_evp2d_dump

  type(c_ptr), save :: handle = c_null_ptr
  logical(kind=log_kind), save :: ldebug = .false., linfo = .true., ldevpack = .true.
  logical(kind=log_kind), save :: lautotune = .false., tuned = .false.
  integer(kind=int_kind), save :: ts = 0, iu06 = -1, buf1d = 0
  logical(kind=log_kind), save :: log_open = .false.   ! newunit numbers are negative
  integer(kind=int_kind), save :: nx, ny, nActive, navel, nghost, nx_block, ny_block, max_block
  integer(kind=int_kind), allocatable, save :: indxti(:), indxtj(:)
  real(kind=dbl_kind), allocatable, save :: buf(:)
  integer(c_int32_t), allocatable, save :: ibuf(:)
  real(kind=dbl_kind), pointer, save :: gall(:,:,:) => null()   ! 2D staging (page-locked)
  type(c_ptr), save :: gall_mem = c_null_ptr
  ! the fields dyn_evp1d_run packs (argument order, then uvel/vvel as _init)
  ! and the fields it unpacks
  integer(c_int32_t), parameter :: fin(31) = [F_STRESSP_1, F_STRESSP_2, F_STRESSP_3, F_STRESSP_4, &
      F_STRESSM_1, F_STRESSM_2, F_STRESSM_3, F_STRESSM_4, F_STRESS12_1, F_STRESS12_2, F_STRESS12_3, &
      F_STRESS12_4, F_STRENGTH, F_CDN_OCN, F_AIU, F_UOCN, F_VOCN, F_WATERX, F_WATERY, F_FORCEX,    &
      F_FORCEY, F_UMASSDTI, F_FM, F_STRINTX, F_STRINTY, F_TBU, F_CB, F_UVEL, F_VVEL,               &
      F_UVEL_INIT, F_VVEL_INIT]
  integer(c_int32_t), parameter :: fout(17) = [F_STRESSP_1, F_STRESSP_2, F_STRESSP_3, F_STRESSP_4, &
      F_STRESSM_1, F_STRESSM_2, F_STRESSM_3, F_STRESSM_4, F_STRESS12_1, F_STRESS12_2, F_STRESS12_3, &
      F_STRESS12_4, F_STRINTX, F_STRINTY, F_CB, F_UVEL, F_VVEL]

contains

  !=============================================================================
  subroutine fill_params(p)
    use ice_dyn_shared, only: ndte, capping, e_factor, epp2i, deltaminEVP, arlx, arlx1i, &
                              brlx, denom1, revp, Ktens, u0, cosw, sinw
    type(evp1d_params_t), intent(out) :: p
    p%ndte = ndte
    p%capping = capping
    p%e_factor = e_factor
    p%epp2i = epp2i
    p%deltaminEVP = deltaminEVP
    p%arlx = arlx
    p%arlx1i = arlx1i
    p%brlx = brlx
    p%denom1 = denom1
    p%revp = revp
    p%Ktens = Ktens
    p%u0 = u0
    p%cosw = cosw
    p%sinw = sinw
    p%rhow = 1026.0_dbl_kind          ! stepu's local parameter (bench_v2b.F90:489-490)
  end subroutine fill_params

  !=============================================================================
  subroutine read_nml()
    integer(kind=int_kind) :: lun, ios
    namelist /evp1d_nml/ ldebug, linfo, buf1d, ldevpack, lautotune
    open(newunit=lun, file='evp_1d.nml', status='old', iostat=ios)
    if (ios == 0) then
      read(lun, nml=evp1d_nml, iostat=ios)
      close(lun)
    endif
    if (buf1d /= 0) error stop 'ice_dyn_evp1d: buf1d /= 0 is not supported (static point list)'
  end subroutine read_nml

  !=============================================================================
  subroutine dyn_evp1d_init(cnx, cny, cnx_block, cny_block, cmax_block, cnghost, &
                            L_dyT, L_dxT, L_uarear, L_tmask, G_HTE, G_HTN)
    use ice_communicate, only: master_task
    use ice_domain, only: distrb_info
    use ice_gather_scatter, only: gather_global_ext
    integer(kind=int_kind), intent(in) :: cnx, cny, cnx_block, cny_block, cmax_block, cnghost
    real(kind=dbl_kind), dimension(:,:,:), intent(in) :: L_dyT, L_dxT, L_uarear
    logical(kind=log_kind), dimension(:,:,:), intent(in) :: L_tmask
    real(kind=dbl_kind), dimension(:,:), intent(in) :: G_HTE, G_HTN
    real(kind=dbl_kind), allocatable :: G_dyT(:,:), G_dxT(:,:), G_uarear(:,:)
    logical(kind=log_kind), allocatable :: G_tmask(:,:)
    integer(c_int32_t), allocatable :: imask(:,:), indxtij(:), nbr(:,:), ti(:), tj(:)
    type(evp1d_params_t) :: p
    integer(c_int32_t) :: na32, navel32
    integer(kind=int_kind) :: k, ios

    if (log_open) close(iu06)
    open(newunit=iu06, file='evp1d.log', status='replace', iostat=ios)
    if (ios /= 0) error stop 'Unable to open evp1d.log'
    log_open = .true.
    write(iu06,'(a)') 'Initializing evp1d (MI355X engine)'
    call read_nml()
    tuned = .false.
    if (c_associated(handle)) call dyn_evp1d_finalize_handle()
    ts = 0
    nx_block = cnx_block
    ny_block = cny_block
    max_block = cmax_block
    nghost = cnghost
    nx = cnx + 2*nghost
    ny = cny + 2*nghost
    allocate(G_dyT(nx,ny), G_dxT(nx,ny), G_uarear(nx,ny), G_tmask(nx,ny), imask(nx,ny))
    call gather_global_ext(G_uarear, L_uarear, master_task, distrb_info)
    call gather_global_ext(G_dxT, L_dxT, master_task, distrb_info)
    call gather_global_ext(G_dyT, L_dyT, master_task, distrb_info)
    call gather_global_ext(G_tmask, L_tmask, master_task, distrb_info)
    imask = merge(1_c_int32_t, 0_c_int32_t, G_tmask)

    ! 1D list: calc_nActiveTU, calc_2d_indices_init, calc_navel, convert_2d_1d_init
    call evp1d_check(evp1d_index_count(int(nx,c_int32_t), int(ny,c_int32_t), int(nghost,c_int32_t), &
                                       imask, na32, navel32), 'evp1d_index_count')
    nActive = na32
    navel = navel32
    if (allocated(indxti)) deallocate(indxti, indxtj, buf, ibuf)
    allocate(indxti(max(nActive,1)), indxtj(max(nActive,1)), indxtij(max(navel,1)), nbr(max(nActive,1),6))
    allocate(ti(max(nActive,1)), tj(max(nActive,1)), buf(max(navel,1)), ibuf(max(nActive,1)))
    call evp1d_check(evp1d_index_build(int(nx,c_int32_t), int(ny,c_int32_t), int(nghost,c_int32_t), imask, &
                     na32, navel32, ti, tj, indxtij, nbr(:,1), nbr(:,2), nbr(:,3), nbr(:,4),         &
                     nbr(:,5), nbr(:,6)), 'evp1d_index_build')
    indxti(1:nActive) = ti(1:nActive)
    indxtj(1:nActive) = tj(1:nActive)
    write(iu06,*) 'nx, ny, nActive, navel: ', nx, ny, nActive, navel

    call fill_params(p)
    call evp1d_check(evp1d_create(na32, navel32, p, handle), 'evp1d_create')
    do k = 1, 6
      call evp1d_check(evp1d_set_index(handle, int(k-1,c_int32_t), nbr(:,k), na32), 'evp1d_set_index')
    enddo
    call evp1d_check(evp1d_set_grid(handle, int(nx,c_int32_t), int(ny,c_int32_t), indxtij, navel32), &
                     'evp1d_set_grid')
    call free_gall()
    if (ldevpack) then
      call evp1d_check(evp1d_host_alloc(int(nx,c_int64_t) * ny * size(fin) * 8, gall_mem), 'evp1d_host_alloc')
      call c_f_pointer(gall_mem, gall, [nx, ny, size(fin)])
    endif
    ! static metrics (convert_2d_1d_init, ice_dyn_evp1d.F90:803-815)
    call put_static(G_uarear, 0, 0, F_UAREAR)
    call put_static(G_dxT, 0, 0, F_DXT)
    call put_static(G_dyT, 0, 0, F_DYT)
    call put_static(G_HTE, 0, 0, F_HTE)
    call put_static(G_HTN, 0, 0, F_HTN)
    call put_static(G_HTE, -1, 0, F_HTEM1)
    call put_static(G_HTN, 0, -1, F_HTNM1)
    deallocate(G_dyT, G_dxT, G_uarear, G_tmask, imask, indxtij, nbr, ti, tj)
  end subroutine dyn_evp1d_init

  subroutine put_static(G, di, dj, field)
    real(kind=dbl_kind), intent(in) :: G(:,:)
    integer, intent(in) :: di, dj
    integer(c_int32_t), intent(in) :: field
    integer(kind=int_kind) :: iw
    do iw = 1, nActive
      buf(iw) = G(indxti(iw) + di, indxtj(iw) + dj)
    enddo
    call evp1d_check(evp1d_set_field(handle, field, buf, int(nActive,c_int32_t)), 'evp1d_set_field')
  end subroutine put_static

  ! 2D -> 1D (convert_2d_1d_dyn, ice_dyn_evp1d.F90:845-881); uvel/vvel also
  ! fill their ghost slots with zero (880-881)
  subroutine put_dyn(G, field, with_ghosts)
    real(kind=dbl_kind), intent(in) :: G(:,:)
    integer(c_int32_t), intent(in) :: field
    logical, intent(in) :: with_ghosts
    integer(kind=int_kind) :: iw, n
    do iw = 1, nActive
      buf(iw) = G(indxti(iw), indxtj(iw))
    enddo
    n = nActive
    if (with_ghosts) then
      buf(nActive+1:navel) = c0
      n = navel
    endif
    call evp1d_check(evp1d_set_field(handle, field, buf, int(n,c_int32_t)), 'evp1d_set_field')
  end subroutine put_dyn

  ! 1D -> 2D (convert_1d_2d_dyn, ice_dyn_evp1d.F90:911-946)
  subroutine get_dyn(G, field, with_ghosts)
    real(kind=dbl_kind), intent(inout) :: G(:,:)
    integer(c_int32_t), intent(in) :: field
    logical, intent(in) :: with_ghosts
    integer(kind=int_kind) :: iw, n
    n = nActive
    if (with_ghosts) n = navel
    call evp1d_check(evp1d_get_field(handle, field, buf, int(n,c_int32_t)), 'evp1d_get_field')
    do iw = 1, nActive
      G(indxti(iw), indxtj(iw)) = buf(iw)
    enddo
  end subroutine get_dyn

  !=============================================================================
  subroutine dyn_evp1d_run(L_stressp_1 , L_stressp_2 , L_stressp_3, L_stressp_4,     &
                           L_stressm_1 , L_stressm_2 , L_stressm_3, L_stressm_4,     &
                           L_stress12_1, L_stress12_2, L_stress12_3,L_stress12_4,    &
                           L_strength,                                               &
                           L_cdn_ocn   , L_aiu       , L_uocn     , L_vocn     ,     &
                           L_waterxU   , L_wateryU   , L_forcexU  , L_forceyU  ,     &
                           L_umassdti  , L_fmU       , L_strintxU , L_strintyU ,     &
                           L_Tbu       , L_Cb        , L_uvel      , L_vvel,         &
                           L_icetmask , L_iceUmask                                   )
    use ice_dyn_shared, only: ndte
    use ice_communicate, only: master_task
    use ice_domain, only: distrb_info
    use ice_gather_scatter, only: gather_global_ext, scatter_global_ext
    real(kind=dbl_kind), dimension(:,:,:), intent(inout) ::                          &
                           L_stressp_1 , L_stressp_2 , L_stressp_3, L_stressp_4,     &
                           L_stressm_1 , L_stressm_2 , L_stressm_3, L_stressm_4,     &
                           L_stress12_1, L_stress12_2, L_stress12_3,L_stress12_4,    &
                           L_strintxU , L_strintyU, L_Cb, L_uvel, L_vvel
    real(kind=dbl_kind), dimension(:,:,:), intent(in) ::                             &
                           L_strength,                                               &
                           L_cdn_ocn   , L_aiu       , L_uocn     , L_vocn     ,     &
                           L_waterxU   , L_wateryU   , L_forcexU  , L_forceyU  ,     &
                           L_umassdti  , L_fmU       , L_Tbu
    logical(kind=log_kind), dimension(:,:,:), intent(in) :: L_iceUmask, L_iceTmask
    real(kind=dbl_kind), allocatable :: G(:,:)
    logical(kind=log_kind), allocatable :: GT(:,:), GU(:,:)
    type(evp1d_params_t) :: p
    real(c_float) :: tune_us
    integer(c_int32_t) :: tw, th, tb, tx, ntl, dep, dact, dblk, hset, hact
    integer(kind=int_kind) :: iw

    if (.not. c_associated(handle)) error stop 'dyn_evp1d_run: dyn_evp1d_init was not called'
    allocate(G(nx,ny), GT(nx,ny), GU(nx,ny))

    ! skip flags (set_skipMe, ice_dyn_evp1d.F90:521-536)
    call gather_global_ext(GT, L_iceTmask, master_task, distrb_info)
    call gather_global_ext(GU, L_iceUmask, master_task, distrb_info)
    do iw = 1, nActive
      ibuf(iw) = merge(0_c_int32_t, 1_c_int32_t, GT(indxti(iw), indxtj(iw)))
    enddo
    call evp1d_check(evp1d_set_skip(handle, S_SKIPT, ibuf, int(nActive,c_int32_t)), 'evp1d_set_skip')
    do iw = 1, nActive
      ibuf(iw) = merge(0_c_int32_t, 1_c_int32_t, GU(indxti(iw), indxtj(iw)))
    enddo
    call evp1d_check(evp1d_set_skip(handle, S_SKIPU, ibuf, int(nActive,c_int32_t)), 'evp1d_set_skip')
    if (ldebug) write(iu06,*) 'time step, active points, ice T, ice U: ', ts, nActive, &
                              count(GT), count(GU)

    if (ldevpack) then
      ! 2D -> 1D on the device; fin's order is the argument order.  Each field
      ! is queued for the device as soon as it is gathered, so its copy runs
      ! while the next one is gathered (gall is page-locked and stays
      ! untouched until evp1d_get_2d below has waited for the stream); the
      ! initial velocities are uvel and vvel's own staging.
      call gather_put_async(1, L_stressp_1);    call gather_put_async(2, L_stressp_2)
      call gather_put_async(3, L_stressp_3);    call gather_put_async(4, L_stressp_4)
      call gather_put_async(5, L_stressm_1);    call gather_put_async(6, L_stressm_2)
      call gather_put_async(7, L_stressm_3);    call gather_put_async(8, L_stressm_4)
      call gather_put_async(9, L_stress12_1);   call gather_put_async(10, L_stress12_2)
      call gather_put_async(11, L_stress12_3);  call gather_put_async(12, L_stress12_4)
      call gather_put_async(13, L_strength);    call gather_put_async(14, L_cdn_ocn)
      call gather_put_async(15, L_aiu);         call gather_put_async(16, L_uocn)
      call gather_put_async(17, L_vocn);        call gather_put_async(18, L_waterxU)
      call gather_put_async(19, L_wateryU);     call gather_put_async(20, L_forcexU)
      call gather_put_async(21, L_forceyU);     call gather_put_async(22, L_umassdti)
      call gather_put_async(23, L_fmU);         call gather_put_async(24, L_strintxU)
      call gather_put_async(25, L_strintyU);    call gather_put_async(26, L_Tbu)
      call gather_put_async(27, L_Cb);          call gather_put_async(28, L_uvel)
      call gather_put_async(29, L_vvel)
      call evp1d_check(evp1d_put_2d_async(handle, 1_c_int32_t, fin(30:30), gall(:,:,28)), 'evp1d_put_2d_async')
      call evp1d_check(evp1d_put_2d_async(handle, 1_c_int32_t, fin(31:31), gall(:,:,29)), 'evp1d_put_2d_async')
    else
      call gather_put(L_stressp_1, F_STRESSP_1);   call gather_put(L_stressp_2, F_STRESSP_2)
      call gather_put(L_stressp_3, F_STRESSP_3);   call gather_put(L_stressp_4, F_STRESSP_4)
      call gather_put(L_stressm_1, F_STRESSM_1);   call gather_put(L_stressm_2, F_STRESSM_2)
      call gather_put(L_stressm_3, F_STRESSM_3);   call gather_put(L_stressm_4, F_STRESSM_4)
      call gather_put(L_stress12_1, F_STRESS12_1); call gather_put(L_stress12_2, F_STRESS12_2)
      call gather_put(L_stress12_3, F_STRESS12_3); call gather_put(L_stress12_4, F_STRESS12_4)
      call gather_put(L_strength, F_STRENGTH)
      call gather_put(L_cdn_ocn, F_CDN_OCN);       call gather_put(L_aiu, F_AIU)
      call gather_put(L_uocn, F_UOCN);             call gather_put(L_vocn, F_VOCN)
      call gather_put(L_waterxU, F_WATERX);        call gather_put(L_wateryU, F_WATERY)
      call gather_put(L_forcexU, F_FORCEX);        call gather_put(L_forceyU, F_FORCEY)
      call gather_put(L_umassdti, F_UMASSDTI);     call gather_put(L_fmU, F_FM)
      call gather_put(L_strintxU, F_STRINTX);      call gather_put(L_strintyU, F_STRINTY)
      call gather_put(L_Tbu, F_TBU);               call gather_put(L_Cb, F_CB)
      call gather_global_ext(G, L_uvel, master_task, distrb_info)
      call put_dyn(G, F_UVEL, .true.)
      call put_dyn(G, F_UVEL_INIT, .false.)
      call gather_global_ext(G, L_vvel, master_task, distrb_info)
      call put_dyn(G, F_VVEL, .true.)
      call put_dyn(G, F_VVEL_INIT, .false.)
    endif

    ! the subcycles (ice_dyn_evp1d.F90:188-207), device resident
    call fill_params(p)
    call evp1d_check(evp1d_set_params(handle, p), 'evp1d_set_params')
    if (lautotune .and. .not. tuned) then
      call evp1d_check(evp1d_autotune_fused(handle, int(min(ndte, 100),c_int32_t), tune_us), 'autotune_fused')
      if (linfo .and. log_open) then
        call evp1d_check(evp1d_get_tiling(handle, tw, th, tb, tx, ntl), 'evp1d_get_tiling')
        call evp1d_check(evp1d_get_fused_depth(handle, dep, dact, dblk), 'evp1d_get_fused_depth')
        call evp1d_check(evp1d_get_cache_hints(handle, hset, hact), 'evp1d_get_cache_hints')
        write(iu06, '(a,f10.2,a,i0,a,i0,a,i0,a,i0,a,i0)') ' evp1d: autotune_fused ', tune_us, &
          ' us/subcycle, tile ', tw, 'x', th, ', workgroup ', tb, ', depth ', merge(2, 1, dact /= 0), &
          ', cache hints ', hact
      endif
      tuned = .true.
    endif
    call evp1d_check(evp1d_run(handle, int(ndte,c_int32_t)), 'evp1d_run')

    if (ldevpack) then
      ! 1D -> 2D on the device into gall(:,:,1:17) (fout is fin's first 12, 24,
      ! 25, 27, 28, 29 in that order), then back to the caller's blocks
      call evp1d_check(evp1d_get_2d(handle, int(size(fout),c_int32_t), fout, gall), 'evp1d_get_2d')
      call scatter_global_ext(L_stressp_1, gall(:,:,1), master_task, distrb_info)
      call scatter_global_ext(L_stressp_2, gall(:,:,2), master_task, distrb_info)
      call scatter_global_ext(L_stressp_3, gall(:,:,3), master_task, distrb_info)
      call scatter_global_ext(L_stressp_4, gall(:,:,4), master_task, distrb_info)
      call scatter_global_ext(L_stressm_1, gall(:,:,5), master_task, distrb_info)
      call scatter_global_ext(L_stressm_2, gall(:,:,6), master_task, distrb_info)
      call scatter_global_ext(L_stressm_3, gall(:,:,7), master_task, distrb_info)
      call scatter_global_ext(L_stressm_4, gall(:,:,8), master_task, distrb_info)
      call scatter_global_ext(L_stress12_1, gall(:,:,9), master_task, distrb_info)
      call scatter_global_ext(L_stress12_2, gall(:,:,10), master_task, distrb_info)
      call scatter_global_ext(L_stress12_3, gall(:,:,11), master_task, distrb_info)
      call scatter_global_ext(L_stress12_4, gall(:,:,12), master_task, distrb_info)
      call scatter_global_ext(L_strintxU, gall(:,:,13), master_task, distrb_info)
      call scatter_global_ext(L_strintyU, gall(:,:,14), master_task, distrb_info)
      call scatter_global_ext(L_Cb, gall(:,:,15), master_task, distrb_info)
      call scatter_global_ext(L_uvel, gall(:,:,16), master_task, distrb_info)
      call scatter_global_ext(L_vvel, gall(:,:,17), master_task, distrb_info)
    else
      call get_scatter(L_stressp_1, F_STRESSP_1);   call get_scatter(L_stressp_2, F_STRESSP_2)
      call get_scatter(L_stressp_3, F_STRESSP_3);   call get_scatter(L_stressp_4, F_STRESSP_4)
      call get_scatter(L_stressm_1, F_STRESSM_1);   call get_scatter(L_stressm_2, F_STRESSM_2)
      call get_scatter(L_stressm_3, F_STRESSM_3);   call get_scatter(L_stressm_4, F_STRESSM_4)
      call get_scatter(L_stress12_1, F_STRESS12_1); call get_scatter(L_stress12_2, F_STRESS12_2)
      call get_scatter(L_stress12_3, F_STRESS12_3); call get_scatter(L_stress12_4, F_STRESS12_4)
      call get_scatter(L_strintxU, F_STRINTX);      call get_scatter(L_strintyU, F_STRINTY)
      call get_scatter(L_Cb, F_CB)
      call get_scatter(L_uvel, F_UVEL, .true.);     call get_scatter(L_vvel, F_VVEL, .true.)
    endif
    deallocate(G, GT, GU)
    ts = ts + 1

  contains

    subroutine gather_put(L, field)
      real(kind=dbl_kind), intent(in) :: L(:,:,:)
      integer(c_int32_t), intent(in) :: field
      call gather_global_ext(G, L, master_task, distrb_info)
      call put_dyn(G, field, .false.)
    end subroutine gather_put

    ! field k of fin: gathered into its page-locked staging gall(:,:,k), then
    ! queued for the device
    subroutine gather_put_async(k, L)
      integer, intent(in) :: k
      real(kind=dbl_kind), intent(in) :: L(:,:,:)
      call gather_global_ext(gall(:,:,k), L, master_task, distrb_info)
      call evp1d_check(evp1d_put_2d_async(handle, 1_c_int32_t, fin(k:k), gall(:,:,k)), 'evp1d_put_2d_async')
    end subroutine gather_put_async

    subroutine get_scatter(L, field, with_ghosts)
      real(kind=dbl_kind), intent(inout) :: L(:,:,:)
      integer(c_int32_t), intent(in) :: field
      logical, intent(in), optional :: with_ghosts
      logical :: wg
      wg = .false.
      if (present(with_ghosts)) wg = with_ghosts
      call gather_global_ext(G, L, master_task, distrb_info)
      call get_dyn(G, field, wg)
      call scatter_global_ext(L, G, master_task, distrb_info)
    end subroutine get_scatter

  end subroutine dyn_evp1d_run

  !=============================================================================
  ! Writes the 29 dynamic 2D fields, after gathering, to <mydebugF>_<ts>.bin
  ! (stream, Fortran order, the argument order of dyn_evp1d_run).
  subroutine dyn_evp2d_dump(L_stressp_1 , L_stressp_2 , L_stressp_3, L_stressp_4,     &
                            L_stressm_1 , L_stressm_2 , L_stressm_3, L_stressm_4,     &
                            L_stress12_1, L_stress12_2, L_stress12_3,L_stress12_4,    &
                            L_strength,                                               &
                            L_cdn_ocn   , L_aiu       , L_uocn     , L_vocn     ,     &
                            L_waterxU   , L_wateryU   , L_forcexU  , L_forceyU  ,     &
                            L_umassdti  , L_fmU       , L_strintxU , L_strintyU ,     &
                            L_Tbu       , L_Cb        , L_uvel     , L_vvel     ,     &
                            L_iceTmask  , L_iceUmask, mydebugF)
    use ice_communicate, only: master_task
    use ice_domain, only: distrb_info
    use ice_gather_scatter, only: gather_global_ext
    real(kind=dbl_kind), dimension(:,:,:), intent(in), target ::                     &
                           L_stressp_1 , L_stressp_2 , L_stressp_3, L_stressp_4,     &
                           L_stressm_1 , L_stressm_2 , L_stressm_3, L_stressm_4,     &
                           L_stress12_1, L_stress12_2, L_stress12_3,L_stress12_4,    &
                           L_strength,                                               &
                           L_cdn_ocn   , L_aiu       , L_uocn     , L_vocn     ,     &
                           L_waterxU   , L_wateryU   , L_forcexU  , L_forceyU  ,     &
                           L_umassdti  , L_fmU       , L_strintxU , L_strintyU ,     &
                           L_Tbu       , L_Cb        , L_uvel     , L_vvel
    logical(kind=log_kind), dimension(:,:,:), intent(in) :: L_iceUmask, L_iceTmask
    character(len=*), intent(in) :: mydebugF
    real(kind=dbl_kind), allocatable :: G(:,:)
    character(len=256) :: fname
    integer(kind=int_kind) :: lun, ios
    allocate(G(nx,ny))
    write(fname,'(a,a,i6.6,a)') trim(mydebugF), '_', ts - 1, '.bin'
    open(newunit=lun, file=trim(fname), form='unformatted', access='stream', status='replace', iostat=ios)
    if (ios /= 0) error stop 'dyn_evp2d_dump: cannot open dump file'
    call w(L_stressp_1);  call w(L_stressp_2);  call w(L_stressp_3);  call w(L_stressp_4)
    call w(L_stressm_1);  call w(L_stressm_2);  call w(L_stressm_3);  call w(L_stressm_4)
    call w(L_stress12_1); call w(L_stress12_2); call w(L_stress12_3); call w(L_stress12_4)
    call w(L_strength);   call w(L_cdn_ocn);    call w(L_aiu);        call w(L_uocn)
    call w(L_vocn);       call w(L_waterxU);    call w(L_wateryU);    call w(L_forcexU)
    call w(L_forceyU);    call w(L_umassdti);   call w(L_fmU);        call w(L_strintxU)
    call w(L_strintyU);   call w(L_Tbu);        call w(L_Cb);         call w(L_uvel)
    call w(L_vvel)
    close(lun)
    deallocate(G)
    if (count(L_iceTmask) < 0 .or. count(L_iceUmask) < 0) error stop 'unreachable'
  contains
    subroutine w(L)
      real(kind=dbl_kind), intent(in) :: L(:,:,:)
      call gather_global_ext(G, L, master_task, distrb_info)
      write(lun) G
    end subroutine w
  end subroutine dyn_evp2d_dump

  !=============================================================================
  subroutine free_gall()
    if (c_associated(gall_mem)) call evp1d_check(evp1d_host_free(gall_mem), 'evp1d_host_free')
    gall_mem = c_null_ptr
    gall => null()
  end subroutine free_gall

  subroutine dyn_evp1d_finalize_handle()
    if (c_associated(handle)) then
      call evp1d_check(evp1d_destroy(handle), 'evp1d_destroy')
      handle = c_null_ptr
    endif
  end subroutine dyn_evp1d_finalize_handle

  subroutine dyn_evp1d_finalize()
    call dyn_evp1d_finalize_handle()
    if (allocated(indxti)) deallocate(indxti, indxtj, buf, ibuf)
    call free_gall()
    write(*,'(a21)') 'Time to die for evp1d'
    if (log_open) then
      write(iu06,'(a21)') 'Time to die for evp1d'
      close(iu06)
      iu06 = -1
      log_open = .false.
    endif
  end subroutine dyn_evp1d_finalize

end module ice_dyn_evp1d
