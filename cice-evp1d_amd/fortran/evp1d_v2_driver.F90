!===============================================================================
! evp1d_v2_driver - the reference's v2 benchmark program on the MI355X engine.
!
! Reads the reference's inputs from the working directory exactly as its
! drivers do (kernel.nml, src/vars.F90:19-69; input_double_1d_v2.bin,
! input_integer_1d.bin, input_logical_1d.bin, src/vars.F90:2101-2135, with the
! four scalars of the double file overriding set_evp_parameters), runs ndte
! subcycles of stress and stepu on the GPU (src/evp_v2b.F90:91-110), prints
! "Benchmark took: <s> seconds" for the subcycle loop alone (the reference's
! 'bench' timer, evp_v2b.F90:84-117) and writes output_stress_v2g_1d.bin and
! output_stepu_v2g_1d.bin in the format of writeout_1d (src/vars.F90:2233-2304).
! lindividual re-times stress-only and stepu-only loops (evp_v2b.F90:128-155).
!
! Unlike the reference, every read is checked (the reference ignores iostat,
! src/vars.F90:2106-2133) and logicals are read as 4-byte integers, nonzero
! meaning .true., so files written by any compiler are accepted.  testscale
! must be 1 (the reference's replicate() is a timing-only device whose copies
! do not line up, src/vars.F90:1801-1838).
!
! Usage:  evp1d_v2_driver [graph|persistent|launch|fused|fused-inplace|fused-pair|tuned]
! (no argument: the library's default, the pair form where the lists tile;
! fused: one subcycle per fused launch; fused-inplace: the same with
! stresses updated in place; fused-pair: two subcycles per fused launch on
! 31x15 tiles; tuned: fused form with the tile, stress storage and depth
! chosen by evp1d_autotune_fused)
!===============================================================================
program evp1d_v2_driver
  use, intrinsic :: iso_c_binding
  use evp1d_capi
  implicit none
  integer, parameter :: dbl = selected_real_kind(13)
  integer(4) :: ndte, nx_block, ny_block, testscale, na, navel, nb, max_blocks, nghost
  logical :: lconvert21d, lindividual, binoutput
  character(len=200) :: inpfname
  namelist /kernel_nml/ inpfname, ndte, nx_block, ny_block, testscale, na, navel, nb, &
                        lconvert21d, lindividual, binoutput, max_blocks, nghost
  real(dbl), allocatable :: f(:,:), uvel(:), vvel(:), ht(:,:), out(:,:), vout(:)
  integer(c_int32_t), allocatable :: idx(:,:), skipU(:), skipT(:)
  real(dbl) :: scal(4), secs
  type(evp1d_params_t) :: p
  type(c_ptr) :: h
  integer :: lun, ios, k, iw, nbb, i
  integer(8) :: t0, t1, rate
  character(len=32) :: arg
  integer(c_int32_t) :: mode, inplace, depth
  logical :: tune, lgrad
  real(dbl) :: loss
  real(dbl), allocatable :: g(:,:)
  real(c_float) :: tune_us
  ! fields of input_double_1d_v2.bin after strength/uvel/vvel, in file order
  integer(c_int32_t), parameter :: mid(28) = [F_DXT, F_DYT, F_UAREAR, F_CDN_OCN, F_AIU, F_UOCN, &
      F_VOCN, F_WATERX, F_WATERY, F_FORCEX, F_FORCEY, F_UMASSDTI, F_FM, F_STRINTX, F_STRINTY,     &
      F_TBU, F_STRESSP_1, F_STRESSP_2, F_STRESSP_3, F_STRESSP_4, F_STRESSM_1, F_STRESSM_2,        &
      F_STRESSM_3, F_STRESSM_4, F_STRESS12_1, F_STRESS12_2, F_STRESS12_3, F_STRESS12_4]
  integer(c_int32_t), parameter :: hid(4) = [F_HTE, F_HTN, F_HTEM1, F_HTNM1]

  ndte = 1; inpfname = 'pathtoinputfile'; nb = -1; na = -1; navel = -1
  nx_block = -1; ny_block = -1; testscale = 1; lconvert21d = .false.
  lindividual = .false.; binoutput = .true.; max_blocks = 1; nghost = 1
  open(newunit=lun, file='kernel.nml', status='old', iostat=ios)
  if (ios /= 0) error stop 'ERROR: kernel.nml not found'
  read(lun, nml=kernel_nml, iostat=ios)
  if (ios /= 0) error stop 'ERROR: Can not read namelist: kernel_nml from file kernel.nml'
  close(lun)
  if (testscale /= 1) error stop 'ERROR: testscale /= 1 is not supported'
  if (na < 0 .or. navel < na) error stop 'ERROR: kernel.nml needs 0 <= na <= navel'
  write(*,'(a,i0,a,i0,a,i0)') ' na = ', na, '  navel = ', navel, '  ndte = ', ndte

  mode = MODE_AUTO; inplace = 0; depth = 0; tune = .false.
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    select case (trim(arg))
    case ('graph');      mode = MODE_GRAPH
    case ('persistent'); mode = MODE_PERSISTENT
    case ('launch');     mode = MODE_LAUNCH
    case ('fused');      mode = MODE_FUSED; depth = 1
    case ('fused-inplace'); mode = MODE_FUSED; inplace = 1; depth = 1
    case ('fused-pair'); mode = MODE_FUSED; depth = 2
    case ('tuned');      mode = MODE_FUSED; tune = .true.
    case default;        error stop 'usage: evp1d_v2_driver [graph|persistent|launch|fused|fused-inplace|fused-pair|tuned] [grad]'
    end select
  endif
  ! grad: the subcycles run as evp1d_adjoint_run (whole trajectory when it
  ! fits), which leaves the same state as evp1d_run, and the gradient of
  ! L = 1/2 sum_{skipU=0} (uvel - uvel_in)^2 + (vvel - vvel_in)^2 with respect
  ! to strength, cdn_ocn, forcex and forcey goes to output_gradient_v2g_1d.bin
  ! (L, then the four arrays, na each, in the input's point order)
  lgrad = .false.
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    if (trim(arg) /= 'grad') error stop 'usage: evp1d_v2_driver [mode] [grad]'
    lgrad = .true.
  endif

  allocate(f(na,28), uvel(navel), vvel(navel), ht(na,4), idx(na,6), skipU(na), skipT(na))
  allocate(out(na,12), vout(max(na,1)))
  open(newunit=lun, file='input_double_1d_v2.bin', form='unformatted', access='stream', &
       action='read', status='old', iostat=ios)
  if (ios /= 0) error stop 'ERROR: cannot open input_double_1d_v2.bin'
  read(lun, iostat=ios) vout(1:na), uvel, vvel, f, scal, ht
  if (ios /= 0) error stop 'ERROR: input_double_1d_v2.bin is short or unreadable'
  close(lun)
  open(newunit=lun, file='input_integer_1d.bin', form='unformatted', access='stream', &
       action='read', status='old', iostat=ios)
  if (ios /= 0) error stop 'ERROR: cannot open input_integer_1d.bin'
  read(lun, iostat=ios) idx
  if (ios /= 0) error stop 'ERROR: input_integer_1d.bin is short or unreadable'
  close(lun)
  open(newunit=lun, file='input_logical_1d.bin', form='unformatted', access='stream', &
       action='read', status='old', iostat=ios)
  if (ios /= 0) error stop 'ERROR: cannot open input_logical_1d.bin'
  read(lun, iostat=ios) skipU, skipT
  if (ios /= 0) error stop 'ERROR: input_logical_1d.bin is short or unreadable'
  close(lun)

  ! set_evp_parameters(dt=300) then the file's scalars (src/vars.F90:2116)
  call evp1d_check(evp1d_set_evp_parameters(int(ndte,c_int32_t), 300.0_dbl, p), 'set_evp_parameters')
  p%capping = scal(1); p%e_factor = scal(2); p%epp2i = scal(3); p%deltaminEVP = scal(4)
  call evp1d_check(evp1d_create(int(na,c_int32_t), int(navel,c_int32_t), p, h), 'evp1d_create')
  call evp1d_check(evp1d_set_mode(h, mode), 'evp1d_set_mode')
  do k = 1, 6
    call evp1d_check(evp1d_set_index(h, int(k-1,c_int32_t), idx(:,k), int(na,c_int32_t)), 'set_index')
  enddo
  call evp1d_check(evp1d_set_skip(h, S_SKIPT, skipT, int(na,c_int32_t)), 'set_skip')
  call evp1d_check(evp1d_set_skip(h, S_SKIPU, skipU, int(na,c_int32_t)), 'set_skip')
  call evp1d_check(evp1d_set_field(h, F_STRENGTH, vout, int(na,c_int32_t)), 'set_field')
  call evp1d_check(evp1d_set_field(h, F_UVEL, uvel, int(navel,c_int32_t)), 'set_field')
  call evp1d_check(evp1d_set_field(h, F_VVEL, vvel, int(navel,c_int32_t)), 'set_field')
  call evp1d_check(evp1d_set_field(h, F_UVEL_INIT, uvel, int(na,c_int32_t)), 'set_field')
  call evp1d_check(evp1d_set_field(h, F_VVEL_INIT, vvel, int(na,c_int32_t)), 'set_field')
  do k = 1, 28
    call evp1d_check(evp1d_set_field(h, mid(k), f(:,k), int(na,c_int32_t)), 'set_field')
  enddo
  do k = 1, 4
    call evp1d_check(evp1d_set_field(h, hid(k), ht(:,k), int(na,c_int32_t)), 'set_field')
  enddo
  call evp1d_check(evp1d_set_fused_inplace(h, inplace), 'set_fused_inplace')
  if (depth > 0) &   ! 0: the library's default (the pair form where the lists tile)
    call evp1d_check(evp1d_set_fused_depth(h, depth), 'set_fused_depth')
  if (depth == 2) &
    call evp1d_check(evp1d_set_tiling(h, 31_c_int32_t, 15_c_int32_t, 512_c_int32_t, 1_c_int32_t), 'set_tiling')
  if (tune) then
    call evp1d_check(evp1d_autotune_fused(h, int(min(ndte, 100),c_int32_t), tune_us), 'autotune_fused')
    write(*,'(a,f10.2,a)') ' autotune_fused: ', tune_us, ' us per subcycle'
  endif
  call evp1d_check(evp1d_synchronize(h), 'synchronize')

  call system_clock(t0, rate)
  if (lgrad) then
    allocate(g(max(na,1),4))
    call evp1d_check(evp1d_adjoint_run(h, int(ndte,c_int32_t), 0_c_int32_t, uvel, vvel, loss, g(:,1), g(:,2), &
                                       g(:,3), g(:,4)), 'evp1d_adjoint_run')
  else
    call evp1d_check(evp1d_run(h, int(ndte,c_int32_t)), 'evp1d_run')
  endif
  call evp1d_check(evp1d_synchronize(h), 'synchronize')
  call system_clock(t1)
  secs = real(t1 - t0, dbl) / real(rate, dbl)
  write(*,'(a,f12.6,a)') 'Benchmark took: ', secs, ' seconds'

  if (binoutput) then
    do k = 1, 12
      call evp1d_check(evp1d_get_field(h, F_STRESSP_1 + int(k-1,c_int32_t), out(:,k), int(na,c_int32_t)), 'get')
    enddo
    open(newunit=lun, file='output_stress_v2g_1d.bin', form='unformatted', access='stream', &
         action='write', status='replace', iostat=ios)
    if (ios /= 0) error stop 'Failed open output_stress_v2g_1d.bin'
    write(lun, iostat=ios) out
    if (ios /= 0) error stop 'Failed write output_stress_v2g_1d.bin'
    close(lun)
    nbb = count(skipU == 0)
    if (nb >= 0 .and. nb /= nbb) write(*,'(a,i0,a,i0)') ' note: nb = ', nb, ' but active U points = ', nbb
    open(newunit=lun, file='output_stepu_v2g_1d.bin', form='unformatted', access='stream', &
         action='write', status='replace', iostat=ios)
    if (ios /= 0) error stop 'Failed open output_stepu_v2g_1d.bin'
    do k = 1, 4
      select case (k)
      case (1); call evp1d_check(evp1d_get_field(h, F_STRINTX, vout, int(na,c_int32_t)), 'get')
      case (2); call evp1d_check(evp1d_get_field(h, F_STRINTY, vout, int(na,c_int32_t)), 'get')
      case (3); call evp1d_check(evp1d_get_field(h, F_UVEL, uvel, int(navel,c_int32_t)), 'get')
                vout(1:na) = uvel(1:na)
      case (4); call evp1d_check(evp1d_get_field(h, F_VVEL, vvel, int(navel,c_int32_t)), 'get')
                vout(1:na) = vvel(1:na)
      end select
      i = 0
      do iw = 1, na
        if (skipU(iw) /= 0) cycle
        i = i + 1
        out(i,1) = vout(iw)
      enddo
      if (i /= nbb) error stop 'PROBLEM in generating output_stepu_v2g_1d.bin'
      if (nbb > 0) write(lun, iostat=ios) out(1:nbb,1)
      if (ios /= 0) error stop 'Failed write output_stepu_v2g_1d.bin'
    enddo
    close(lun)
  endif
  if (lgrad) then
    open(newunit=lun, file='output_gradient_v2g_1d.bin', form='unformatted', access='stream', &
         action='write', status='replace', iostat=ios)
    if (ios /= 0) error stop 'Failed open output_gradient_v2g_1d.bin'
    write(lun, iostat=ios) loss, g(1:na,:)
    if (ios /= 0) error stop 'Failed write output_gradient_v2g_1d.bin'
    close(lun)
    write(*,'(a,es24.16)') ' gradient: loss = ', loss
  endif
  write(*,'(a,i0,1x,i0,1x,i0)') ' Kernel dim (v2g - MI355X HIP): ', na, navel, ndte

  if (lindividual) then
    call system_clock(t0)
    call evp1d_check(evp1d_run_individual(h, 0_c_int32_t, int(ndte,c_int32_t)), 'run_individual')
    call evp1d_check(evp1d_synchronize(h), 'synchronize')
    call system_clock(t1)
    write(*,'(a,f12.6,a)') 'Benchmark took stress: ', real(t1 - t0, dbl) / real(rate, dbl), ' seconds'
    call system_clock(t0)
    call evp1d_check(evp1d_run_individual(h, 1_c_int32_t, int(ndte,c_int32_t)), 'run_individual')
    call evp1d_check(evp1d_synchronize(h), 'synchronize')
    call system_clock(t1)
    write(*,'(a,f12.6,a)') 'Benchmark took stepu: ', real(t1 - t0, dbl) / real(rate, dbl), ' seconds'
  endif
  call evp1d_check(evp1d_destroy(h), 'evp1d_destroy')
end program evp1d_v2_driver
