! Drop-in replacement of ADPRES's hot-path modules `cmfd` and `nodal`
! (src/mod_cmfd.f90, src/mod_nodal.f90).
!
! It exports the names the unchanged host modules use -- outer, outer_ad,
! PowDis (mod_control, mod_th), outer_fs, outer_th, outer_tr, Integrate
! (mod_control, mod_th, mod_trans) and nodal::Lxyz (mod_trans) -- and runs
! the static outer iteration on the GPU through the C-ABI.  The solver state
! stays in the engine's context between calls (the reference keeps it in
! SAVEd module variables), and the results are copied back into the host's
! sdata arrays f0, fs0 and Ke after every call, so the host's outputs
! (AsmPow, AxiPow, AsmFlux) and callers (BCSEARCH) see what they expect.
! Printed lines follow the reference's formats (src/mod_cmfd.f90:331,
! 367-374, 491-494, 500-504, 583-584, 589-591; src/mod_nodal.f90:131-142).
!
! The fixed-source outer (outer_fs) runs on the GPU with k held fixed, the
! T-H outer (outer_th) as a bounded forward outer between T-H updates, and
! the transient outer (outer_tr) with the time step's external source formed
! here on the host and the transient nodal update on the device.


module cmfd

  use iso_c_binding
  use sdata, only: dp
  use adpres_gpu
  implicit none
  save
  private
  public :: outer, outer_ad, outer_fs, outer_th, outer_tr, PowDis, Integrate, leakage_of

  type(c_ptr) :: ctx = c_null_ptr
  logical     :: started = .false.
  logical     :: fixed_mode = .false.   ! outer_fs prints no k
  real(c_double), allocatable :: leak(:,:,:)   ! Lxyz of the device state (nnod, ng, 3)
  logical     :: leak_ok = .false.
  integer(c_int), allocatable, target :: gix(:), giy(:), giz(:)
  real(c_double), allocatable, target :: gxd(:), gyd(:), gzd(:)
  real(c_double), allocatable, target :: tk(:), tser(:), tfer(:), tnd(:)
  integer(c_int), allocatable, target :: tfl(:), tloc(:)

contains

  subroutine ensure_context()
    use sdata, only: nnod, ng, nmat, nxx, nyy, nzz, ix, iy, iz, xdel, ydel, zdel, &
                     xeast, xwest, ynorth, ysouth, zbott, ztop
    type(adg_geom_t) :: geo
    integer(c_int) :: rc
    character(len=16) :: pstr
    if (c_associated(ctx)) return
    allocate(gix(nnod), giy(nnod), giz(nnod), gxd(nxx), gyd(nyy), gzd(nzz))
    gix = ix; giy = iy; giz = iz
    gxd = xdel; gyd = ydel; gzd = zdel
    geo%nnod = nnod; geo%ng = ng; geo%nmat = nmat
    geo%nxx = nxx; geo%nyy = nyy; geo%nzz = nzz
    geo%ix = c_loc(gix); geo%iy = c_loc(giy); geo%iz = c_loc(giz)
    geo%xdel = c_loc(gxd); geo%ydel = c_loc(gyd); geo%zdel = c_loc(gzd)
    geo%bc = [xeast, xwest, ynorth, ysouth, zbott, ztop]
    geo%precision = ADG_FP64
    call get_environment_variable('ADPRES_GPU_PRECISION', pstr)
    if (trim(pstr) == 'FP32') geo%precision = ADG_FP32_REFINE
    rc = adg_create(geo, ctx)
    if (rc /= ADG_OK) then
      write(*,*) '  GPU ENGINE INITIALISATION FAILED, STATUS ', rc
      stop
    end if
    rc = adg_set_profiling(ctx, 1_c_int)
  end subroutine ensure_context

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc == ADG_OK) return
    write(*,*) '  ', what, ' FAILED: ', trim(c_string(adg_last_error(ctx)))
    stop
  end subroutine

  ! one solve on the device (shared by outer, outer_ad, outer_fs, outer_th
  ! and outer_tr; maxn: an outer limit reached without a stop, reported in
  ! hit when present)
  subroutine run_outer(popt, mode, label, maxn, hit)
    use sdata, only: nnod, ng, nout, nin, serc, ferc, nac, nupd, kern, f0, fs0, Ke, &
                     D, sigr, sigs, nuf, sigf, chi, mat, dc, exsrc, ndmax, im, jm, km, &
                     fdm_time, nod_time, ser, fer
    use io, only: ounit, scr
    integer, intent(in) :: popt
    integer(c_int), intent(in) :: mode
    character(len=*), intent(in) :: label
    integer, optional, intent(in) :: maxn
    logical, optional, intent(out) :: hit
    type(adg_iter_t)  :: it
    type(adg_trace_t) :: tr
    integer(c_int)    :: rc, rco, first
    integer :: p, nlast
    real(c_double), target :: kk
    real(c_double), allocatable, target :: rf0(:,:), rfs0(:)
    real(c_double) :: ms(4)
    integer(c_int), allocatable :: imat(:)

    call ensure_context()
    fixed_mode = mode == ADG_FIXED_SOURCE
    allocate(imat(nnod))
    imat = mat
    call check(adg_set_xs(ctx, D, sigr, sigs, nuf, sigf, chi, imat, dc, exsrc), 'ADG_SET_XS')
    first = 0
    if (.not. started) then
      if (.not. allocated(f0)) allocate(f0(nnod, ng))
      if (.not. allocated(fs0)) allocate(fs0(nnod))
      call check(adg_set_state(ctx, c_null_ptr, c_null_ptr, 1._c_double, c_null_ptr, 1_c_int), 'ADG_SET_STATE')
      started = .true.
    end if

    it%nout = nout; it%nin = nin; it%nac = nac; it%nupd = nupd
    if (present(maxn)) it%nout = maxn
    it%serc = serc; it%ferc = ferc
    it%kern = ADG_KERN_SANM
    if (kern == ' FDM') it%kern = ADG_KERN_FDM
    if (kern == ' PNM') it%kern = ADG_KERN_PNM

    if (allocated(tk)) deallocate(tk, tser, tfer, tnd, tfl, tloc)
    allocate(tk(it%nout), tser(it%nout), tfer(it%nout), tnd(it%nout), tfl(it%nout), tloc(3 * it%nout))
    tr%cap = it%nout; tr%n = 0
    tr%k = c_loc(tk); tr%ser = c_loc(tser); tr%fer = c_loc(tfer); tr%ndmax = c_loc(tnd)
    tr%flags = c_loc(tfl); tr%loc = c_loc(tloc)

    rco = adg_outer(ctx, mode, it, tr)

    nlast = tr%n
    if (rco == ADG_E_NODAL_UNSTABLE .or. rco == ADG_E_LU_PIVOT) nlast = tr%n - 1
    do p = 1, nlast
      call print_iteration(popt, p, tr%n)
    end do

    allocate(rf0(nnod, ng), rfs0(nnod))
    rc = adg_get_state(ctx, c_loc(rf0), c_loc(rfs0), c_loc(kk), c_null_ptr, c_null_ptr)
    f0 = rf0
    fs0 = rfs0
    Ke = kk
    if (tr%n > 0) then
      ndmax = tnd(tr%n); im = tloc(3*tr%n-2); jm = tloc(3*tr%n-1); km = tloc(3*tr%n)
      ser = tser(tr%n); fer = tfer(tr%n)   ! the last outer's errors, as the callers print them
    end if
    leak_ok = .false.
    call update_s0(mode)
    rc = adg_last_timing(ctx, ms)
    fdm_time = fdm_time + (ms(1) - ms(3)) * 1.e-3_dp
    nod_time = nod_time + ms(3) * 1.e-3_dp

    if (present(hit)) hit = rco == ADG_E_MAX_OUTER
    if (present(maxn) .and. rco == ADG_E_MAX_OUTER) rco = ADG_OK
    select case (rco)
    case (ADG_OK)
    case (ADG_E_MAX_OUTER)
      write(*,*)
      write(*,*) '  MAXIMUM NUMBER OF OUTER ITERATION IS REACHED IN ', label, ' CALCULATION.'
      write(*,*) '  CHECK PROBLEM SPECIFICATION OR CHANGE ITERATION CONTROL (%ITER).'
      write(*,*) '  PERHAPS BY MAKING FISSION SOURCE INTERPOLATION MORE FREQUENT'
      write(*,*) '  ADPRES IS STOPING...'
      stop
    case (ADG_E_NODAL_UNSTABLE)
      if (popt > 0 .and. iand(tfl(tr%n), 1) /= 0) call say(' ', '    ...FISSION SOURCE EXTRAPOLATED...')
      write(*,*)
      write(*,'(" Error: Max. change in nodal coupling coefficient = ", F10.1)') tnd(tr%n)
      write(*,*)"The two-node nonlinear iteration seems not stable."
      write(*,*)"Try to reduce time step size for transient..."
      write(*,*)"and/or change iteration control using %ITER card, ..."
      write(*,*)"perhaps by making nodal update less frequent."
      write(*,*)"Or increase number inner iteration per outer iteration."
      stop
    case (ADG_E_LU_PIVOT)
      write(ounit,*) 'ERROR IN MATRIX DECOMP: DIAGONAL ELEMENTS CLOSE TO ZERO'
      write(*,*) 'ERROR IN MATRIX DECOMP: DIAGONAL ELEMENTS CLOSE TO ZERO'
      write(*,*) '  ', trim(c_string(adg_last_error(ctx)))
      stop
    case default
      call check(rco, 'ADG_OUTER')
    end select

    if (popt > 0 .and. .not. fixed_mode) then
      write(ounit,*)
      write(ounit,'(2X,"MULTIPLICATION EFFECTIVE (K-EFF) = ", F9.6)') Ke
      if (scr) then
        write(*,*)
        write(*,'(2X,"MULTIPLICATION EFFECTIVE (K-EFF) = ", F9.6)') Ke
      end if
    end if
  end subroutine run_outer

  subroutine say(dummy, line)
    use io, only: ounit, scr
    character(len=*), intent(in) :: dummy, line
    write(ounit,*) line
    if (scr) write(*,*) line
    if (.false.) write(*,*) dummy
  end subroutine

  subroutine print_iteration(popt, p, n)
    use io, only: ounit, scr
    integer, intent(in) :: popt, p, n
    if (popt <= 0 .or. p > n) return
    if (iand(tfl(p), 1) /= 0) call say(' ', '    ...FISSION SOURCE EXTRAPOLATED...')
    if (iand(tfl(p), 2) /= 0) then
      call say(' ', '    .....NODAL COUPLING UPDATED..... ')
      write(ounit,1145) tnd(p), tloc(3*p-2), tloc(3*p-1), tloc(3*p)
      if (scr) write(*,1145) tnd(p), tloc(3*p-2), tloc(3*p-1), tloc(3*p)
    end if
    if (fixed_mode) then
      write(ounit,'(I5,2ES15.5)') p, tser(p), tfer(p)
      if (scr) write(*,'(I5,2ES15.5)') p, tser(p), tfer(p)
    else
      write(ounit,'(I5,F13.6,2ES15.5)') p, tk(p), tser(p), tfer(p)
      if (scr) write(*,'(I5,F13.6,2ES15.5)') p, tk(p), tser(p), tfer(p)
    end if
1145 FORMAT ('MAX. CHANGE IN NODAL COUPLING COEF.= ', ES12.5, &
             ' AT NODE I = ', I2, ', J = ', I2, ', K = ', I2)
  end subroutine

  subroutine outer(popt)
    integer, optional, intent(in) :: popt
    integer :: po
    po = 0
    if (present(popt)) po = popt
    call run_outer(po, ADG_FORWARD, 'FORWARD')
  end subroutine outer

  subroutine outer_ad(popt)
    integer, optional, intent(in) :: popt
    integer :: po
    po = 0
    if (present(popt)) po = popt
    ! with popt = 0 (rod_eject) the reference keeps the forward D-hats
    ! (src/mod_cmfd.f90:677-679)
    if (po > 0) then
      call run_outer(po, ADG_ADJOINT, 'ADJOINT')
    else
      call run_outer(po, ADG_ADJOINT_FROZEN, 'ADJOINT')
    end if
  end subroutine outer_ad

  subroutine outer_fs(popt)
    integer, optional, intent(in) :: popt
    integer :: po
    po = 0
    if (present(popt)) po = popt
    call run_outer(po, ADG_FIXED_SOURCE, 'FIXED-SOURCE')
  end subroutine outer_fs

  ! outer_th (src/mod_cmfd.f90:703-796): at most maxn outers of the forward
  ! iteration between two T-H updates, printing nothing; without an %ITER
  ! card the nodal update interval becomes nth / 2
  subroutine outer_th(maxn)
    use sdata, only: nth, nupd
    use io, only: biter
    integer, intent(in) :: maxn
    if (biter == 0) nupd = int(nth/2)
    call run_outer(0, ADG_FORWARD, 'T-H', maxn)
  end subroutine outer_th

  ! outer_tr (src/mod_cmfd.f90:800-880): the transient fixed-source outer of
  ! one time step ht, after the step's external source; maxi when nout
  ! outers did not converge
  subroutine outer_tr(ht, maxi)
    use sdata, only: nnod, nout, tbeta, mat, dfis
    real(dp), intent(in) :: ht
    logical, intent(out) :: maxi
    real(c_double), allocatable :: w(:)
    integer :: n
    call transient_source(ht)
    allocate(w(nnod))
    do n = 1, nnod
      w(n) = 1._dp - tbeta(mat(n)) + dfis(n)
    end do
    call ensure_context()
    call check(adg_set_transient(ctx, w), 'ADG_SET_TRANSIENT')
    call run_outer(0, ADG_TRANSIENT, 'TRANSIENT', nout, maxi)
  end subroutine outer_tr

  ! The transient's external source and delayed-fission weight of a time
  ! step ht, into sdata's exsrc and dfis (the reference's get_exsrc,
  ! src/mod_cmfd.f90:884-952): precursors decayed over the step, the
  ! previous step's flux and its theta-method terms (bth, sth).  Without
  ! branch tables the decay factors depend on the precursor group only and
  ! are formed once per group.
  subroutine transient_source(ht)
    use sdata, only: nnod, ng, nf, lamb, c0, iBeta, chi, mat, velo, fst, ft, omeg, dfis, nuf, &
                     bth, sth, s0, tbeta, sigrp, L, m, exsrc
    use io, only: bxtab
    real(dp), intent(in) :: ht
    real(dp), allocatable :: pxg(:), a1g(:), a2g(:)
    real(dp) :: pxe, a1, a2, dt, dtp, pthet, lam, ib, ve
    integer :: n, i, g
    allocate(pxg(nf), a1g(nf), a2g(nf))
    if (bxtab /= 1) then
      do i = 1, nf
        pxg(i) = exp(-lamb(i)*ht)
        a1g(i) = (1._dp - pxg(i)) / (lamb(i)*ht)
        a2g(i) = 1._dp - a1g(i)
        a1g(i) = a1g(i) - pxg(i)
      end do
    end if
    dfis = 0._dp
    do n = 1, nnod
      dt = 0._dp; dtp = 0._dp
      do i = 1, nf
        if (bxtab == 1) then
          lam = m(mat(n))%lamb(i); ib = m(mat(n))%iBeta(i)
          pxe = exp(-lam*ht)
          if (nuf(n,ng) > 0.) then
            a1 = (1._dp - pxe) / (lam*ht)
          else
            a1 = 0._dp
          end if
          a2 = 1._dp - a1
          a1 = a1 - pxe
        else
          lam = lamb(i); ib = iBeta(i)
          pxe = pxg(i); a1 = a1g(i); a2 = a2g(i)
        end if
        dfis(n) = dfis(n) + ib * a2
        dt  = dt + lam * c0(n,i) * pxe + ib * a1 * fst(n)
        dtp = dtp + lam * c0(n,i)
      end do
      do g = 1, ng
        if (bxtab == 1) then
          ve = m(mat(n))%velo(g)
        else
          ve = velo(g)
        end if
        pthet = -L(n,g) - sigrp(n,g) * ft(n,g) + s0(n,g) &
              + (1._dp - tbeta(mat(n))) * chi(mat(n),g) * fst(n) + chi(mat(n),g) * dtp
        exsrc(n,g) = chi(mat(n),g) * dt + exp(omeg(n,g) * ht) * ft(n,g) / (sth * ve * ht) + bth * pthet
      end do
    end do
  end subroutine transient_source

  ! sdata's s0 as the reference leaves it after a solve: TSrc, TSrcAd and
  ! TSrcTr (src/mod_cmfd.f90:1006-1096) clear all of s0 and fill the group
  ! being solved, so after the sweep it holds the in-scatter source of the
  ! group solved last (the last group; the first in the adjoint sweep) from
  ! the final flux.  The transient source reads it.
  subroutine update_s0(mode)
    use sdata, only: nnod, ng, sigs, f0, s0
    integer(c_int), intent(in) :: mode
    integer :: g, h, n
    if (.not. allocated(s0)) allocate(s0(nnod, ng))
    s0 = 0._dp
    g = ng
    if (mode == ADG_ADJOINT .or. mode == ADG_ADJOINT_FROZEN) g = 1
    do h = 1, ng
      if (h == g) cycle
      do n = 1, nnod
        if (mode == ADG_ADJOINT .or. mode == ADG_ADJOINT_FROZEN) then
          s0(n,g) = s0(n,g) + sigs(n,g,h) * f0(n,h)
        else
          s0(n,g) = s0(n,g) + sigs(n,h,g) * f0(n,h)
        end if
      end do
    end do
  end subroutine update_s0

  ! Lxyz of node n, group g of the device state, fetched once per solve
  subroutine leakage_of(n, g, L1, L2, L3)
    use sdata, only: nnod, ng
    integer, intent(in)   :: n, g
    real(dp), intent(out) :: L1, L2, L3
    if (.not. leak_ok) then
      if (.not. allocated(leak)) allocate(leak(nnod, ng, 3))
      call check(adg_leakage(ctx, leak), 'ADG_LEAKAGE')
      leak_ok = .true.
    end if
    L1 = leak(n,g,1); L2 = leak(n,g,2); L3 = leak(n,g,3)
  end subroutine leakage_of

  ! node powers from the device (PowDis, src/mod_cmfd.f90:1290-1333)
  subroutine PowDis(p)
    use io, only: ounit
    real(dp), dimension(:), intent(out) :: p
    integer(c_int) :: rc
    rc = adg_powdis(ctx, p)
    if (rc == ADG_E_POWER_NONPOS) then
      write(ounit, *) '   ERROR: TOTAL NODES POWER IS ZERO OR LESS'
      write(ounit, *) '   STOP IN subroutine POWDIS'
      stop
    end if
    call check(rc, 'ADG_POWDIS')
  end subroutine PowDis

  ! Integrate (src/mod_cmfd.f90:1120-1139): volume integral over the nodes
  function Integrate(s) result(intg)
    use sdata, only: nnod, vdel
    real(dp), dimension(:), intent(in) :: s
    real(dp) :: intg
    integer :: n
    intg = 0._dp
    do n = 1, nnod
      intg = intg + vdel(n) * s(n)
    end do
  end function Integrate

end module cmfd

! nodal::Lxyz for the host's transient routines (mod_trans's reactivity):
! the zeroth transverse leakages of the device state
module nodal
  use sdata, only: dp
  implicit none
contains
  subroutine Lxyz(n, g, L1, L2, L3)
    use cmfd, only: leakage_of
    integer, intent(in)   :: n, g
    real(dp), intent(out) :: L1, L2, L3
    call leakage_of(n, g, L1, L2, L3)
  end subroutine
end module nodal
