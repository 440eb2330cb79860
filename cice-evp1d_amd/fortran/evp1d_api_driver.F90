!===============================================================================
! evp1d_api_driver - a CICE-like host of the evp1d driver API.
!
! Plays the part of the reference's integration/ice_dyn_evp1d_standalone.F90
! (ciceinit, dyn_evp1d_init, dyn_evp1d_run x nrun, dyn_evp1d_finalize) on a
! 2D state read from a stream file, and writes the 29 dynamic 2D fields after
! the last run, so the result can be checked:
!
!   api_in.bin  : int32 nx_global, ny_global, nghost, ndte, nrun
!                 int32 tmask, icetmask, iceumask      (NX x NY each)
!                 f64   dxT, dyT, uarear, HTE, HTN     (NX x NY each)
!                 f64   the 29 dynamic fields in dyn_evp1d_run argument order
!   api_out.bin : f64   the 29 dynamic fields after nrun calls of dyn_evp1d_run
! with NX = nx_global + 2*nghost, NY likewise, one block (max_block = 1).
! Prints the wall time of every dyn_evp1d_run call (2D<->1D data movement and
! the ndte subcycles).
!
! Usage:  evp1d_api_driver [input [output]]
!===============================================================================
program evp1d_api_driver
  use ice_kinds_mod
  use ice_dyn_shared, only: ndte, set_evp_parameters
  use ice_dyn_evp1d
  implicit none
  integer(4) :: hdr(5)
  integer(kind=int_kind) :: nxg, nyg, ng, NX, NY, nrun, r, lun, ios, k
  integer(8) :: t0, t1, rate
  integer(4), allocatable :: im(:,:,:)
  real(kind=dbl_kind), allocatable :: st(:,:,:), dyn(:,:,:,:), G_HTE(:,:), G_HTN(:,:)
  logical(kind=log_kind), allocatable :: tmask(:,:,:), icet(:,:,:), iceu(:,:,:)
  character(len=256) :: fin, fout

  fin = 'api_in.bin'
  fout = 'api_out.bin'
  if (command_argument_count() >= 1) call get_command_argument(1, fin)
  if (command_argument_count() >= 2) call get_command_argument(2, fout)

  open(newunit=lun, file=trim(fin), form='unformatted', access='stream', action='read', &
       status='old', iostat=ios)
  if (ios /= 0) error stop 'evp1d_api_driver: cannot open input'
  read(lun, iostat=ios) hdr
  if (ios /= 0) error stop 'evp1d_api_driver: short header'
  nxg = hdr(1); nyg = hdr(2); ng = hdr(3); ndte = hdr(4); nrun = hdr(5)
  NX = nxg + 2*ng
  NY = nyg + 2*ng
  allocate(im(NX,NY,3), st(NX,NY,5), dyn(NX,NY,1,29), G_HTE(NX,NY), G_HTN(NX,NY))
  allocate(tmask(NX,NY,1), icet(NX,NY,1), iceu(NX,NY,1))
  read(lun, iostat=ios) im, st, dyn
  if (ios /= 0) error stop 'evp1d_api_driver: short input'
  close(lun)
  tmask(:,:,1) = im(:,:,1) /= 0
  icet(:,:,1) = im(:,:,2) /= 0
  iceu(:,:,1) = im(:,:,3) /= 0
  G_HTE = st(:,:,4)
  G_HTN = st(:,:,5)

  call set_evp_parameters(3600.0_dbl_kind)
  call dyn_evp1d_init(nxg, nyg, NX, NY, 1, ng,                                   &
                      st(:,:,2:2), st(:,:,1:1), st(:,:,3:3), tmask, G_HTE, G_HTN)
  do r = 1, nrun
    call system_clock(t0, rate)
    call dyn_evp1d_run(dyn(:,:,:,1), dyn(:,:,:,2), dyn(:,:,:,3), dyn(:,:,:,4),    &
                       dyn(:,:,:,5), dyn(:,:,:,6), dyn(:,:,:,7), dyn(:,:,:,8),    &
                       dyn(:,:,:,9), dyn(:,:,:,10), dyn(:,:,:,11), dyn(:,:,:,12), &
                       dyn(:,:,:,13),                                             &
                       dyn(:,:,:,14), dyn(:,:,:,15), dyn(:,:,:,16), dyn(:,:,:,17), &
                       dyn(:,:,:,18), dyn(:,:,:,19), dyn(:,:,:,20), dyn(:,:,:,21), &
                       dyn(:,:,:,22), dyn(:,:,:,23), dyn(:,:,:,24), dyn(:,:,:,25), &
                       dyn(:,:,:,26), dyn(:,:,:,27), dyn(:,:,:,28), dyn(:,:,:,29), &
                       icet, iceu)
    call system_clock(t1)
    write(*,'(a,i0,a,f10.3,a)') ' dyn_evp1d_run ', r, ': ', 1.0e3_dbl_kind * real(t1 - t0, dbl_kind) / &
                                real(rate, dbl_kind), ' ms'
  enddo
  call dyn_evp2d_dump(dyn(:,:,:,1), dyn(:,:,:,2), dyn(:,:,:,3), dyn(:,:,:,4),     &
                      dyn(:,:,:,5), dyn(:,:,:,6), dyn(:,:,:,7), dyn(:,:,:,8),     &
                      dyn(:,:,:,9), dyn(:,:,:,10), dyn(:,:,:,11), dyn(:,:,:,12),  &
                      dyn(:,:,:,13),                                              &
                      dyn(:,:,:,14), dyn(:,:,:,15), dyn(:,:,:,16), dyn(:,:,:,17), &
                      dyn(:,:,:,18), dyn(:,:,:,19), dyn(:,:,:,20), dyn(:,:,:,21), &
                      dyn(:,:,:,22), dyn(:,:,:,23), dyn(:,:,:,24), dyn(:,:,:,25), &
                      dyn(:,:,:,26), dyn(:,:,:,27), dyn(:,:,:,28), dyn(:,:,:,29), &
                      icet, iceu, 'evp2d_dump')
  call dyn_evp1d_finalize()

  open(newunit=lun, file=trim(fout), form='unformatted', access='stream', action='write', &
       status='replace', iostat=ios)
  if (ios /= 0) error stop 'evp1d_api_driver: cannot open output'
  do k = 1, 29
    write(lun) dyn(:,:,1,k)
  enddo
  close(lun)
end program evp1d_api_driver
