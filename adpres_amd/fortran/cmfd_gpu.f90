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
  public :: outer, outer_ad, outer_fs, outer_th, outer_tr, PowDis, Integrate, leakage_of, dev_xs_update

  type(c_ptr) :: ctx = c_null_ptr
  logical     :: started = .false.
  logical     :: fixed_mode = .false.   ! outer_fs prints no k
  real(c_double), allocatable :: leak(:,:,:)   ! Lxyz of the device state (nnod, ng, 3)
  logical     :: leak_ok = .false.
  integer(c_int), allocatable, target :: gix(:), giy(:), giz(:)
  real(c_double), allocatable, target :: gxd(:), gyd(:), gzd(:)
  real(c_double), allocatable, target :: tk(:), tser(:), tfer(:), tnd(:)
  integer(c_int), allocatable, target :: tfl(:), tloc(:)
  ! node cross sections made on the device (dev_xs_update, behind the
  ! xsec module's XS_updt / XStab_updt): the tables are installed once, and
  ! xs_dev says the device holds the node cross sections of the last
  ! update, so the next outer uploads nothing
  logical :: xs_dev = .false.
  logical :: tabs_in = .false.
  logical :: host_xs_filled = .false.
  real(c_double), allocatable, target :: tbuf(:)        ! feedback tables, or branch points and records
  integer(c_int), allocatable, target :: tdims(:), tfb(:), tmat(:)

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
    ! per-phase events only on request: they turn the replayed outer graph
    ! into direct launches; fdm_time / nod_time come from adg_last_timing
    ! either way
    call get_environment_variable('ADPRES_GPU_PROFILE', pstr)
    if (trim(pstr) == '1') rc = adg_set_profiling(ctx, 1_c_int)
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
    ! the device's own cross sections of the last XS_updt need no upload
    if (.not. xs_dev) call check(adg_set_xs(ctx, D, sigr, sigs, nuf, sigf, chi, imat, dc, exsrc), 'ADG_SET_XS')
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

  ! xsec::XS_updt (kind 0, src/mod_xsec.f90:11-46) and xsec::XStab_updt
  ! (kind 1, :48-84) on the device: the tables (feedback tables, or the
  ! branch tables of a %XTAB deck) are installed on the first call -- and
  ! on every call of a transient, whose host changes them (KNE1 scaling of
  ! xnuf, src/mod_trans.f90:509-510) -- and adg_update_xs makes the node
  ! cross sections from the parameters.  The host's node arrays D, sigr,
  ! nuf, sigf, sigs and dc are filled from the device on the first call and
  ! in the modes whose host reads or changes them (transient: sigr gets the
  ! time terms; fixed source: the external source goes with the upload);
  ! otherwise they are left as they were and the next outer uploads nothing.
  subroutine dev_xs_update(kind, xbcon, xftem, xmtem, xcden, xbpos)
    use sdata, only: nnod, ng, dc, D, sigr, nuf, sigf, sigs, mode, xs_time
    use io, only: ounit, bftem, bmtem, bcden, bcrod
    integer, intent(in) :: kind
    real(dp), intent(in) :: xbcon
    real(dp), dimension(:), intent(in) :: xftem, xmtem, xcden, xbpos
    real(c_double), allocatable, target :: pf(:), pm(:), pc(:), pb(:)
    real(c_double), allocatable, target :: gD(:,:), gr(:,:), gn(:,:), gf(:,:), gs(:,:,:), gdc(:,:,:)
    real(c_double) :: ms(4)
    type(c_ptr) :: af, am, ac, ab
    integer(c_int) :: rc
    logical :: host
    call ensure_context()
    host = (.not. host_xs_filled) .or. mode == 'RODEJECT' .or. mode == 'FIXEDSRC'
    if (.not. tabs_in .or. mode == 'RODEJECT') then
      if (kind == 0) then
        call install_feedback_tables()
      else
        call install_branch_tables()
      end if
      tabs_in = .true.
    end if
    ! the parameters of the active feedbacks (a branch-table deck reads the
    ! three node-wise ones always, src/mod_xsec.f90:70-73)
    af = c_null_ptr; am = c_null_ptr; ac = c_null_ptr; ab = c_null_ptr
    if (kind == 1 .or. bftem == 1) then
      allocate(pf(nnod)); pf = xftem(1:nnod); af = c_loc(pf)
    end if
    if (kind == 1 .or. bmtem == 1) then
      allocate(pm(nnod)); pm = xmtem(1:nnod); am = c_loc(pm)
    end if
    if (kind == 1 .or. bcden == 1) then
      allocate(pc(nnod)); pc = xcden(1:nnod); ac = c_loc(pc)
    end if
    if (bcrod == 1) then
      allocate(pb(size(xbpos))); pb = xbpos; ab = c_loc(pb)
    end if
    rc = adg_update_xs(ctx, real(xbcon, c_double), af, am, ac, ab)
    if (rc == ADG_E_XS) then
      write(*,*)
      write(*,*) '  ERROR IN THE CROSS SECTIONS: ', trim(c_string(adg_last_error(ctx)))
      write(ounit,*) '  ERROR IN THE CROSS SECTIONS: ', trim(c_string(adg_last_error(ctx)))
      stop
    end if
    call check(rc, 'ADG_UPDATE_XS')
    rc = adg_last_timing(ctx, ms)
    xs_time = xs_time + ms(1) * 1.e-3_dp
    if (host) then
      allocate(gD(nnod, ng), gr(nnod, ng), gn(nnod, ng), gf(nnod, ng), gs(nnod, ng, ng), gdc(nnod, ng, 6))
      call check(adg_get_xs(ctx, c_loc(gD), c_loc(gr), c_loc(gn), c_loc(gf), c_loc(gs), c_loc(gdc)), 'ADG_GET_XS')
      D = gD; sigr = gr; nuf = gn; sigf = gf; sigs = gs
      if (kind == 1) dc = gdc
      host_xs_filled = .true.
    end if
    xs_dev = .not. (mode == 'RODEJECT' .or. mode == 'FIXEDSRC')
  end subroutine dev_xs_update

  ! sdata's feedback tables (base, boron, fuel and moderator temperature,
  ! coolant density, rods) into one buffer, with pointers per kind and
  ! cross section as adg_xs_tables takes them; a kind whose card is absent
  ! gets no pointers
  subroutine install_feedback_tables()
    use sdata, only: nnod, ng, nmat, nxx, nyy, mat, chi, dc, nb, rbcon, rftem, rmtem, rcden, coreh, pos0, &
                     ssize, fbmap, xsigtr, xsiga, xnuf, xsigf, xsigs, csigtr, csiga, cnuf, csigf, csigs, &
                     fsigtr, fsiga, fnuf, fsigf, fsigs, msigtr, msiga, mnuf, msigf, msigs, &
                     lsigtr, lsiga, lnuf, lsigf, lsigs, dsigtr, dsiga, dnuf, dsigf, dsigs
    use io, only: bbcon, bcbcs, bftem, bmtem, bcden, bcrod
    type(adg_xs_tables_t) :: t
    integer :: kd, q, off, n2, n3
    logical :: on(6)
    n2 = nmat * ng
    n3 = nmat * ng * ng
    on = [.true., bbcon == 1 .or. bcbcs == 1, bftem == 1, bmtem == 1, bcden == 1, bcrod == 1]
    if (allocated(tbuf)) deallocate(tbuf)
    allocate(tbuf(6 * (4 * n2 + n3)))
    tbuf = 0._c_double
    t%nmat = nmat
    t%nb = 0
    if (bcrod == 1) t%nb = nb
    t%rbcon = rbcon; t%rftem = rftem; t%rmtem = rmtem; t%rcden = rcden
    t%coreh = coreh; t%pos0 = pos0; t%ssize = ssize
    t%tab = c_null_ptr
    t%fbmap = c_null_ptr
    off = 0
    do kd = 1, 6
      if (on(kd)) then
        select case (kd)
        case (1)
          call put5(xsigtr, xsiga, xnuf, xsigf, xsigs)
        case (2)
          call put5(csigtr, csiga, cnuf, csigf, csigs)
        case (3)
          call put5(fsigtr, fsiga, fnuf, fsigf, fsigs)
        case (4)
          call put5(msigtr, msiga, mnuf, msigf, msigs)
        case (5)
          call put5(lsigtr, lsiga, lnuf, lsigf, lsigs)
        case (6)
          call put5(dsigtr, dsiga, dnuf, dsigf, dsigs)
        end select
        do q = 1, 5
          t%tab(q, kd) = c_loc(tbuf(off + 1 + (q - 1) * n2))
        end do
      end if
      off = off + 4 * n2 + n3
    end do
    if (bcrod == 1) then
      if (allocated(tfb)) deallocate(tfb)
      allocate(tfb(nxx * nyy))
      tfb = reshape(fbmap(1:nxx, 1:nyy), [nxx * nyy])
      t%fbmap = c_loc(tfb)
    end if
    if (allocated(tmat)) deallocate(tmat)
    allocate(tmat(nnod))
    tmat = mat
    call check(adg_set_xs_tables(ctx, t, tmat, chi, dc), 'ADG_SET_XS_TABLES')
  contains
    subroutine put5(a, b, c, d, e)
      real(dp), intent(in) :: a(:,:), b(:,:), c(:,:), d(:,:), e(:,:,:)
      tbuf(off + 1:off + n2) = reshape(a(1:nmat, 1:ng), [n2])
      tbuf(off + n2 + 1:off + 2 * n2) = reshape(b(1:nmat, 1:ng), [n2])
      tbuf(off + 2 * n2 + 1:off + 3 * n2) = reshape(c(1:nmat, 1:ng), [n2])
      tbuf(off + 3 * n2 + 1:off + 4 * n2) = reshape(d(1:nmat, 1:ng), [n2])
      tbuf(off + 4 * n2 + 1:off + 4 * n2 + n3) = reshape(e(1:nmat, 1:ng, 1:ng), [n3])
    end subroutine
  end subroutine install_feedback_tables

  ! sdata's branch tables m(:) of a %XTAB deck, flattened as
  ! adg_set_xtab_tables takes them: per material (nd, nb, nf, nm, trod),
  ! its branch points, and its records (density fastest, then boron, fuel,
  ! moderator temperature) of sigtr, siga, nuf, sigf, sigs, dc in storage
  ! order; the rodded records of the materials with a rodded set after them
  subroutine install_branch_tables()
    use sdata, only: nnod, ng, nmat, nxx, nyy, mat, chi, nb, coreh, pos0, ssize, fbmap, m
    use io, only: bcrod
    type(adg_xtab_tables_t) :: t
    real(c_double), allocatable, target :: rbuf(:)
    integer :: mm, s, tt, q, v, E, npt, nrec, nrrec, o, orr
    E = 4 * ng + ng * ng + 6 * ng
    npt = 0; nrec = 0; nrrec = 0
    do mm = 1, nmat
      npt = npt + m(mm)%nd + m(mm)%nb + m(mm)%nf + m(mm)%nm
      nrec = nrec + m(mm)%nd * m(mm)%nb * m(mm)%nf * m(mm)%nm
      if (m(mm)%trod == 1) nrrec = nrrec + m(mm)%nd * m(mm)%nb * m(mm)%nf * m(mm)%nm
    end do
    if (allocated(tbuf)) deallocate(tbuf)
    if (allocated(tdims)) deallocate(tdims)
    allocate(tbuf(npt + nrec * E), tdims(5 * nmat), rbuf(max(1, nrrec * E)))
    o = 0
    do mm = 1, nmat
      tdims(5 * (mm - 1) + 1:5 * mm) = [m(mm)%nd, m(mm)%nb, m(mm)%nf, m(mm)%nm, m(mm)%trod]
      tbuf(o + 1:o + m(mm)%nd) = m(mm)%pd(1:m(mm)%nd); o = o + m(mm)%nd
      tbuf(o + 1:o + m(mm)%nb) = m(mm)%pb(1:m(mm)%nb); o = o + m(mm)%nb
      tbuf(o + 1:o + m(mm)%nf) = m(mm)%pf(1:m(mm)%nf); o = o + m(mm)%nf
      tbuf(o + 1:o + m(mm)%nm) = m(mm)%pm(1:m(mm)%nm); o = o + m(mm)%nm
    end do
    orr = 0
    do mm = 1, nmat
      do v = 1, m(mm)%nm
        do q = 1, m(mm)%nf
          do tt = 1, m(mm)%nb
            do s = 1, m(mm)%nd
              call put_rec(tbuf, o, m(mm)%xsec(s, tt, q, v))
              if (m(mm)%trod == 1) call put_rec(rbuf, orr, m(mm)%rxsec(s, tt, q, v))
            end do
          end do
        end do
      end do
    end do
    t%nmat = nmat
    t%nb = 0
    if (bcrod == 1) t%nb = nb
    t%dims = c_loc(tdims)
    t%points = c_loc(tbuf(1))
    t%xsec = c_loc(tbuf(npt + 1))
    t%rxsec = c_null_ptr
    if (nrrec > 0) t%rxsec = c_loc(rbuf)
    t%coreh = coreh; t%pos0 = pos0; t%ssize = ssize
    t%fbmap = c_null_ptr
    if (bcrod == 1) then
      if (allocated(tfb)) deallocate(tfb)
      allocate(tfb(nxx * nyy))
      tfb = reshape(fbmap(1:nxx, 1:nyy), [nxx * nyy])
      t%fbmap = c_loc(tfb)
    end if
    if (allocated(tmat)) deallocate(tmat)
    allocate(tmat(nnod))
    tmat = mat
    call check(adg_set_xtab_tables(ctx, t, tmat, chi), 'ADG_SET_XTAB_TABLES')
  contains
    subroutine put_rec(buf, off, x)
      use sdata, only: XBRANCH
      real(c_double), intent(inout) :: buf(:)
      integer, intent(inout) :: off
      type(XBRANCH), intent(in) :: x
      buf(off + 1:off + ng) = x%sigtr(1:ng); off = off + ng
      buf(off + 1:off + ng) = x%siga(1:ng); off = off + ng
      buf(off + 1:off + ng) = x%nuf(1:ng); off = off + ng
      buf(off + 1:off + ng) = x%sigf(1:ng); off = off + ng
      buf(off + 1:off + ng * ng) = reshape(x%sigs, [ng * ng]); off = off + ng * ng
      buf(off + 1:off + 6 * ng) = reshape(x%dc, [6 * ng]); off = off + 6 * ng
    end subroutine
  end subroutine install_branch_tables

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
