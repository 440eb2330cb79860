!===============================================================================
! evp1d_v0_driver - the reference's v0 (2D block) benchmark on the MI355X engine.
!
! Reads the reference's v0 inputs from the working directory as its driver
! does: kernel.nml (src/vars.F90:19-69), input_double_2d.bin and
! input_integer_2d.bin (readin_2d, src/vars.F90:295-316) and, with
! max_blocks > 1, the halo tables base_output_integer_blockindices_ref_v0_2d.bin
! and base_output_integer_halo_ref_v0_2d.bin (readin_halo, 80-99).  The three
! scalars of the double file override set_evp_parameters(dt=300)
! (src/evp_v0.F90:68-70).  Runs ndte subcycles of stress, stepu and the halo
! update on the GPU (src/evp_v0.F90:93-136), prints "Benchmark took: <s>
! seconds" for that loop alone and writes output_stress_v0_2d.bin and
! output_stepu_v0_2d.bin in the format of writeout_2d (317-353), plus with
! lconvert21d the 1D files of convert_2d_1d (534-584).  lindividual re-times
! stress-only and stepu-only loops (evp_v0.F90:157-204).
!
! Unlike the reference, every read is checked (the reference ignores iostat,
! src/vars.F90:298-314) and testscale must be 1 (resize_2d is a timing-only
! device).
!
! Usage:  evp1d_v0_driver [fused|plain]   (the subcycle form, default fused)
!===============================================================================
program evp1d_v0_driver
  use, intrinsic :: iso_c_binding
  use evp1d_capi
  implicit none
  integer, parameter :: dbl = selected_real_kind(13)
  integer(4) :: ndte, nx_block, ny_block, testscale, na, navel, nb, max_blocks, nghost
  logical :: lconvert21d, lindividual, binoutput
  character(len=200) :: inpfname
  namelist /kernel_nml/ inpfname, ndte, nx_block, ny_block, testscale, na, navel, nb, &
                        lconvert21d, lindividual, binoutput, max_blocks, nghost
  real(dbl), allocatable :: f(:,:), out(:,:), v1(:)
  integer(c_int32_t), allocatable :: ti(:,:), tj(:,:), ui(:,:), uj(:,:), icellt(:), icellu(:)
  integer(c_int32_t), allocatable :: ilo(:), jlo(:), ihi(:), jhi(:), src(:,:), dst(:,:)
  integer(c_int32_t) :: navel_file, nlb, ncp
  real(dbl) :: scal(3), secs
  type(evp1d_params_t) :: p
  type(c_ptr) :: h
  integer :: lun, ios, k, iblk, ij, i, j, ncell
  integer(8) :: t0, t1, rate, n
  character(len=32) :: arg
  integer(c_int32_t) :: mode
  integer, parameter :: nfile = 38   ! arrays of input_double_2d.bin

  mode = MODE2D_FUSED
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    select case (trim(arg))
    case ('fused'); mode = MODE2D_FUSED
    case ('plain'); mode = MODE2D_PLAIN
    case default;   error stop 'usage: evp1d_v0_driver [fused|plain]'
    end select
  endif
  ndte = 1; inpfname = 'pathtoinputfile'; nb = -1; na = -1; navel = -1
  nx_block = -1; ny_block = -1; testscale = 1; lconvert21d = .false.
  lindividual = .false.; binoutput = .true.; max_blocks = 1; nghost = 1
  open(newunit=lun, file='kernel.nml', status='old', iostat=ios)
  if (ios /= 0) error stop 'ERROR: kernel.nml not found'
  read(lun, nml=kernel_nml, iostat=ios)
  if (ios /= 0) error stop 'ERROR: Can not read namelist: kernel_nml from file kernel.nml'
  close(lun)
  if (testscale /= 1) error stop 'ERROR: testscale /= 1 is not supported'
  if (nx_block < 2 .or. ny_block < 2 .or. max_blocks < 1) &
    error stop 'ERROR: kernel.nml needs nx_block, ny_block >= 2 and max_blocks >= 1'
  ncell = nx_block * ny_block
  n = int(ncell, 8) * max_blocks
  write(*,'(a,i0,a,i0,a,i0,a,i0)') ' nx_block = ', nx_block, '  ny_block = ', ny_block, &
       '  max_blocks = ', max_blocks, '  ndte = ', ndte

  allocate(f(n, nfile), out(n, 4))
  open(newunit=lun, file='input_double_2d.bin', form='unformatted', access='stream', &
       action='read', status='old', iostat=ios)
  if (ios /= 0) error stop 'ERROR: cannot open input_double_2d.bin'
  read(lun, iostat=ios) f, scal
  if (ios /= 0) error stop 'ERROR: input_double_2d.bin is short or unreadable'
  close(lun)
  allocate(ti(ncell, max_blocks), tj(ncell, max_blocks), ui(ncell, max_blocks), uj(ncell, max_blocks))
  allocate(icellt(max_blocks), icellu(max_blocks))
  open(newunit=lun, file='input_integer_2d.bin', form='unformatted', access='stream', &
       action='read', status='old', iostat=ios)
  if (ios /= 0) error stop 'ERROR: cannot open input_integer_2d.bin'
  read(lun, iostat=ios) ti, tj, ui, uj, icellt, icellu, navel_file
  if (ios /= 0) error stop 'ERROR: input_integer_2d.bin is short or unreadable'
  close(lun)

  ! set_evp_parameters(dt=300), then the file's scalars (src/vars.F90:305)
  call evp1d_check(evp1d_set_evp_parameters(int(ndte,c_int32_t), 300.0_dbl, p), 'set_evp_parameters')
  p%capping = scal(1); p%e_factor = scal(2); p%epp2i = scal(3)
  call evp1d_check(evp2d_create(int(nx_block,c_int32_t), int(ny_block,c_int32_t), &
                                int(max_blocks,c_int32_t), p, h), 'evp2d_create')
  call evp1d_check(evp2d_set_mode(h, mode), 'evp2d_set_mode')
  do k = 1, nfile
    call evp1d_check(evp2d_set_field(h, int(k-1,c_int32_t), f(:,k), n), 'evp2d_set_field')
  enddo
  ! uvel_init = uvel, vvel_init = vvel (readin_2d, src/vars.F90:307-308)
  call evp1d_check(evp2d_set_field(h, V_UVEL_INIT, f(:,V_UVEL+1), n), 'evp2d_set_field')
  call evp1d_check(evp2d_set_field(h, V_VVEL_INIT, f(:,V_VVEL+1), n), 'evp2d_set_field')
  do iblk = 1, max_blocks
    call evp1d_check(evp2d_set_cells(h, int(iblk,c_int32_t), icellt(iblk), ti(:,iblk), tj(:,iblk), &
                                     icellu(iblk), ui(:,iblk), uj(:,iblk)), 'evp2d_set_cells')
  enddo
  if (max_blocks > 1) then
    open(newunit=lun, file='base_output_integer_blockindices_ref_v0_2d.bin', form='unformatted', &
         access='stream', action='read', status='old', iostat=ios)
    if (ios /= 0) error stop 'ERROR: cannot open base_output_integer_blockindices_ref_v0_2d.bin'
    read(lun, iostat=ios) nlb
    if (ios /= 0 .or. nlb < 0) error stop 'ERROR: bad base_output_integer_blockindices_ref_v0_2d.bin'
    allocate(ilo(max(nlb,1)), jlo(max(nlb,1)), ihi(max(nlb,1)), jhi(max(nlb,1)))
    read(lun, iostat=ios) ilo(1:nlb), jlo(1:nlb), ihi(1:nlb), jhi(1:nlb)
    if (ios /= 0) error stop 'ERROR: base_output_integer_blockindices_ref_v0_2d.bin is short'
    close(lun)
    open(newunit=lun, file='base_output_integer_halo_ref_v0_2d.bin', form='unformatted', &
         access='stream', action='read', status='old', iostat=ios)
    if (ios /= 0) error stop 'ERROR: cannot open base_output_integer_halo_ref_v0_2d.bin'
    read(lun, iostat=ios) ncp
    if (ios /= 0 .or. ncp < 0) error stop 'ERROR: bad base_output_integer_halo_ref_v0_2d.bin'
    allocate(src(3, max(ncp,1)), dst(3, max(ncp,1)))
    read(lun, iostat=ios) src(:,1:ncp), dst(:,1:ncp)
    if (ios /= 0) error stop 'ERROR: base_output_integer_halo_ref_v0_2d.bin is short'
    close(lun)
    call evp1d_check(evp2d_set_halo(h, nlb, ilo, jlo, ihi, jhi, ncp, src, dst), 'evp2d_set_halo')
  endif
  call evp1d_check(evp2d_synchronize(h), 'synchronize')

  call system_clock(t0, rate)
  call evp1d_check(evp2d_run(h, int(ndte,c_int32_t)), 'evp2d_run')
  call evp1d_check(evp2d_synchronize(h), 'synchronize')
  call system_clock(t1)
  secs = real(t1 - t0, dbl) / real(rate, dbl)
  write(*,'(a,f12.6,a)') 'Benchmark took: ', secs, ' seconds'

  if (binoutput) then
    ! writeout_2d (src/vars.F90:317-334): the 12 stresses, then strintx,
    ! strinty, uvel, vvel, whole arrays
    open(newunit=lun, file='output_stress_v0_2d.bin', form='unformatted', access='stream', &
         action='write', status='replace', iostat=ios)
    if (ios /= 0) error stop 'Failed open output_stress_v0_2d.bin'
    do k = 0, 11
      call evp1d_check(evp2d_get_field(h, V_STRESSP_1 + int(k,c_int32_t), f(:,1), n), 'evp2d_get_field')
      write(lun, iostat=ios) f(:,1)
      if (ios /= 0) error stop 'Failed write output_stress_v0_2d.bin'
    enddo
    close(lun)
    call evp1d_check(evp2d_get_field(h, V_STRINTX, out(:,1), n), 'evp2d_get_field')
    call evp1d_check(evp2d_get_field(h, V_STRINTY, out(:,2), n), 'evp2d_get_field')
    call evp1d_check(evp2d_get_field(h, V_UVEL, out(:,3), n), 'evp2d_get_field')
    call evp1d_check(evp2d_get_field(h, V_VVEL, out(:,4), n), 'evp2d_get_field')
    open(newunit=lun, file='output_stepu_v0_2d.bin', form='unformatted', access='stream', &
         action='write', status='replace', iostat=ios)
    if (ios /= 0) error stop 'Failed open output_stepu_v0_2d.bin'
    write(lun, iostat=ios) out
    if (ios /= 0) error stop 'Failed write output_stepu_v0_2d.bin'
    close(lun)
    if (lconvert21d) then
      ! convert_2d_1d (src/vars.F90:534-584): block 1's T cells for the
      ! stresses, its U cells for strintx, strinty, uvel, vvel
      allocate(v1(max(icellt(1), icellu(1), 1)))
      open(newunit=lun, file='output_stress_v0_1d.bin', form='unformatted', access='stream', &
           action='write', status='replace', iostat=ios)
      if (ios /= 0) error stop 'Failed open output_stress_v0_1d.bin'
      do k = 0, 11
        call evp1d_check(evp2d_get_field(h, V_STRESSP_1 + int(k,c_int32_t), f(:,1), n), 'evp2d_get_field')
        do ij = 1, icellt(1)
          i = ti(ij,1); j = tj(ij,1)
          v1(ij) = f(i + (j-1)*nx_block, 1)
        enddo
        if (icellt(1) > 0) write(lun, iostat=ios) v1(1:icellt(1))
        if (ios /= 0) error stop 'Failed write output_stress_v0_1d.bin'
      enddo
      close(lun)
      open(newunit=lun, file='output_stepu_v0_1d.bin', form='unformatted', access='stream', &
           action='write', status='replace', iostat=ios)
      if (ios /= 0) error stop 'Failed open output_stepu_v0_1d.bin'
      do k = 1, 4
        do ij = 1, icellu(1)
          i = ui(ij,1); j = uj(ij,1)
          v1(ij) = out(i + (j-1)*nx_block, k)
        enddo
        if (icellu(1) > 0) write(lun, iostat=ios) v1(1:icellu(1))
        if (ios /= 0) error stop 'Failed write output_stepu_v0_1d.bin'
      enddo
      close(lun)
      write(*,'(a,i0,1x,i0,1x,i0)') ' Kernel 1D parameters na, navel, nb: ', icellt(1), navel_file, icellu(1)
    endif
  endif
  write(*,'(a,i0,1x,i0,1x,i0,1x,i0)') ' Kernel dim (v0 - MI355X HIP): ', sum(icellt), sum(icellu), ndte, max_blocks

  if (lindividual) then
    call system_clock(t0)
    call evp1d_check(evp2d_run_individual(h, 0_c_int32_t, int(ndte,c_int32_t)), 'run_individual')
    call evp1d_check(evp2d_synchronize(h), 'synchronize')
    call system_clock(t1)
    write(*,'(a,f12.6,a)') 'Benchmark took stress: ', real(t1 - t0, dbl) / real(rate, dbl), ' seconds'
    call system_clock(t0)
    call evp1d_check(evp2d_run_individual(h, 1_c_int32_t, int(ndte,c_int32_t)), 'run_individual')
    call evp1d_check(evp2d_synchronize(h), 'synchronize')
    call system_clock(t1)
    write(*,'(a,f12.6,a)') 'Benchmark took stepu: ', real(t1 - t0, dbl) / real(rate, dbl), ' seconds'
  endif
  call evp1d_check(evp2d_destroy(h), 'evp2d_destroy')
end program evp1d_v0_driver
