!===============================================================================
! Single-task stand-ins for the CICE modules that ice_dyn_evp1d uses, for
! building the driver API outside CICE (tests, standalone programs).  Inside
! CICE these files are not compiled: CICE's own modules of the same names are
! used.  Mirrors the roles of integration/ice_kinds_mod.F90,
! ice_constants.F90, ice_dyn_shared.F90, ice_gather_scatter.F90,
! ice_communicate.F90 and ice_domain.F90 of the reference.
!===============================================================================
module ice_kinds_mod
  implicit none
  public
  integer, parameter :: char_len = 80, char_len_long = 256,   &
                        log_kind = kind(.true.),               &
                        int_kind = selected_int_kind(6),       &
                        real_kind = selected_real_kind(6),     &
                        dbl_kind = selected_real_kind(13)
end module ice_kinds_mod

module ice_constants
  use ice_kinds_mod
  implicit none
  public
  real(kind=dbl_kind), parameter :: c0 = 0.0_dbl_kind, c1 = 1.0_dbl_kind, c2 = 2.0_dbl_kind
end module ice_constants

! Parameters of the EVP solver as CICE's ice_dyn_shared holds them
! (values of src/ice.F90:80-130).
module ice_dyn_shared
  use ice_kinds_mod
  use ice_constants
  implicit none
  public
  real(kind=dbl_kind), parameter :: u0 = 5e-5_dbl_kind, cosw = c1, sinw = c0, &
                                    elasticDamp = 0.36_dbl_kind, Ktens = c0, revp = c0, &
                                    e_plasticpot = c2, e_yieldcurve = c2
  integer(kind=int_kind) :: ndte = 1
  real(kind=dbl_kind) :: e_factor, capping, epp2i, brlx, arlx, arlx1i, denom1
  real(kind=dbl_kind) :: deltaminEVP = 1.0e-8_dbl_kind
contains
  subroutine set_evp_parameters(dt)
    real(kind=dbl_kind), intent(in) :: dt   ! unused, as in the reference
    epp2i = c1 / e_plasticpot**2
    capping = c1
    e_factor = e_yieldcurve**2 / e_plasticpot**4
    arlx = c2 * elasticDamp * real(ndte, kind=dbl_kind)
    arlx1i = c1 / arlx
    brlx = real(ndte, kind=dbl_kind)
    denom1 = c1 / (c1 + arlx1i)
    if (dt < c0) stop 'set_evp_parameters: negative dt'
  end subroutine set_evp_parameters
end module ice_dyn_shared

module ice_communicate
  use ice_kinds_mod
  implicit none
  integer(kind=int_kind), parameter :: master_task = 0, my_task = 0
end module ice_communicate

module ice_domain
  use ice_kinds_mod
  implicit none
  integer(kind=int_kind), parameter :: distrb_info = 1
end module ice_domain

! one block, one task: gather copies block 1, scatter copies back
module ice_gather_scatter
  use ice_kinds_mod
  implicit none
  private
  public :: gather_global_ext, scatter_global_ext
  interface gather_global_ext
    module procedure gather_dbl, gather_log
  end interface
  interface scatter_global_ext
    module procedure scatter_dbl
  end interface
contains
  subroutine gather_dbl(ARRAY_G, ARRAY, dst_task, src_dist)
    real(kind=dbl_kind), intent(inout) :: ARRAY_G(:,:)
    real(kind=dbl_kind), intent(in) :: ARRAY(:,:,:)
    integer(kind=int_kind), intent(in) :: dst_task, src_dist
    ARRAY_G = ARRAY(:,:,1)
  end subroutine gather_dbl
  subroutine gather_log(ARRAY_G, ARRAY, dst_task, src_dist)
    logical(kind=log_kind), intent(inout) :: ARRAY_G(:,:)
    logical(kind=log_kind), intent(in) :: ARRAY(:,:,:)
    integer(kind=int_kind), intent(in) :: dst_task, src_dist
    ARRAY_G = ARRAY(:,:,1)
  end subroutine gather_log
  subroutine scatter_dbl(ARRAY, ARRAY_G, src_task, dst_dist)
    real(kind=dbl_kind), intent(inout) :: ARRAY(:,:,:)
    real(kind=dbl_kind), intent(in) :: ARRAY_G(:,:)
    integer(kind=int_kind), intent(in) :: src_task, dst_dist
    ARRAY(:,:,1) = ARRAY_G
  end subroutine scatter_dbl
end module ice_gather_scatter
