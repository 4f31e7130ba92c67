! Drop-in replacement of ADPRES's module xsec (src/mod_xsec.f90) for the
! GPU build: the host modules call XS_updt and XStab_updt (mod_control,
! mod_th, mod_trans) with the T-H parameters and bank positions, and the
! node cross sections are made on the device (cmfd's dev_xs_update:
! adg_set_xs_tables / adg_set_xtab_tables once, then adg_update_xs), bit
! for bit the reference's, so the outer that follows uploads nothing.

module xsec

  use sdata, only: dp
  implicit none
  private
  public :: XS_updt, XStab_updt

contains

  ! xsec::XS_updt (src/mod_xsec.f90:11-46): feedback tables
  subroutine XS_updt(xbcon, xftem, xmtem, xcden, xbpos)
    use cmfd, only: dev_xs_update
    real(dp), intent(in) :: xbcon
    real(dp), dimension(:), intent(in) :: xftem, xmtem, xcden, xbpos
    call dev_xs_update(0, xbcon, xftem, xmtem, xcden, xbpos)
  end subroutine XS_updt

  ! xsec::XStab_updt (src/mod_xsec.f90:48-84): branch tables of a %XTAB deck
  subroutine XStab_updt(xbcon, xftem, xmtem, xcden, xbpos)
    use cmfd, only: dev_xs_update
    real(dp), intent(in) :: xbcon
    real(dp), dimension(:), intent(in) :: xftem, xmtem, xcden, xbpos
    call dev_xs_update(1, xbcon, xftem, xmtem, xcden, xbpos)
  end subroutine XStab_updt

end module xsec
