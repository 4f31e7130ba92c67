! Repo-owned dump driver for the ADPRES reference build (test infrastructure).
!
! Links the reference modules unchanged (sdata, io, xsec, cmfd, nodal) and
! exposes three actions so that the reference's own FP64 arithmetic can be
! written out at full precision:
!
!   ref_dump DECK solve OUT [XSFILE|-]          whole static solve (outer,
!                                               outer_ad or outer_fs by %MODE)
!   ref_dump DECK nodal OUT XSFILE|- STATEFILE  one nodal (SANM/PNM) update
!   ref_dump DECK bicg  OUT XSFILE|- STATEFILE  one nin-iteration BiCGSTAB per group
!   ref_dump DECK trstep OUT XSFILE|-           steady state, then one transient
!                                               step (outer_tr) of a %EJCT deck
!   ref_dump DECK xsupd OUT -                    the cross-section tables and
!                                               XS_updt at the deck's and at
!                                               perturbed T-H / rod parameters
!   ref_dump DECK xtab OUT -                     the same for a %XTAB deck: the
!                                               branch tables and XStab_updt
!
! The deck is read by the reference's own parser (io::inp_read, mod_io.f90:69),
! cross sections are built by xsec::XS_updt (mod_xsec.f90:11).  An optional
! XSFILE overrides the per-node sigtr, siga, nuf, sigf, sigs (raw little-endian
! float64, Fortran order) before xsec::Dsigr_updt (mod_xsec.f90:199) re-derives
! D and sigr; this is how per-node perturbations that no deck can express are
! given to the reference.
!
! OUT is a raw stream of the problem (geometry and node cross sections as the
! reference holds them) followed by the action's results.  The layout is read
! by adpres_amd/problem.py (read_ref_dump).

program ref_dump

  use sdata
  use io,    only: inp_read, bbcon, bcbcs, bftem, bmtem, bcden, bcrod, bxtab
  use xsec,  only: XS_updt, XStab_updt, Dsigr_updt
  use cmfd,  only: outer, outer_ad, outer_fs, outer_tr, PowDis, matrix_setup, bicg
  use nodal, only: nodal_update, nodal_update_pnm

  implicit none

  character(len=256) :: action, outname, xsname, stname
  integer :: u, n, g, nargs, kid, nin_p, j, mm, s, t, q, v
  real(dp) :: kes, ht
  logical :: maxi
  real(dp), allocatable :: f0s(:,:), fs0s(:), dns(:,:,:)
  real(dp), allocatable :: npw(:), bb(:,:), xx(:)
  real(dp), allocatable :: dnin(:,:,:), dfout(:,:,:), dnout(:,:,:)

  nargs = command_argument_count()
  if (nargs < 3) then
    write(*,*) 'usage: ref_dump DECK solve|nodal|bicg OUT [XSFILE|-] [STATEFILE]'
    stop 2
  end if
  call get_command_argument(2, action)
  call get_command_argument(3, outname)
  xsname = '-'
  stname = '-'
  if (nargs >= 4) call get_command_argument(4, xsname)
  if (nargs >= 5) call get_command_argument(5, stname)

  call inp_read()
  if (bxtab == 1) then
    call XStab_updt(bcon, ftem, mtem, cden, bpos)
  else
    call XS_updt(bcon, ftem, mtem, cden, bpos)
  end if

  if (trim(xsname) /= '-') then
    open(newunit=u, file=trim(xsname), access='stream', form='unformatted', &
         status='old', action='read')
    read(u) sigtr, siga, nuf, sigf, sigs
    close(u)
    call Dsigr_updt()
  end if

  kid = 2
  if (kern == ' FDM') kid = 0
  if (kern == ' PNM') kid = 1

  open(newunit=u, file=trim(outname), access='stream', form='unformatted', &
       status='replace', action='write')
  ! ---- problem section ----
  write(u) int(20240601, 4), int(nnod, 4), int(ng, 4), int(nxx, 4), int(nyy, 4), &
           int(nzz, 4), int(nmat, 4), int(kid, 4)
  write(u) int(xeast, 4), int(xwest, 4), int(ynorth, 4), int(ysouth, 4), &
           int(zbott, 4), int(ztop, 4)
  write(u) int(nout, 4), int(nin, 4), int(nac, 4), int(nupd, 4), serc, ferc
  write(u) xdel, ydel, zdel
  write(u) (int(ystag(n)%smin, 4), n = 1, nyy), (int(ystag(n)%smax, 4), n = 1, nyy)
  write(u) (int(xstag(n)%smin, 4), n = 1, nxx), (int(xstag(n)%smax, 4), n = 1, nxx)
  write(u) (int(ix(n), 4), n = 1, nnod), (int(iy(n), 4), n = 1, nnod), &
           (int(iz(n), 4), n = 1, nnod), (int(mat(n), 4), n = 1, nnod)
  write(u) sigtr, siga, nuf, sigf, sigs, chi, dc, D, sigr, vdel

  select case (trim(action))

  case ('solve')
    if (trim(mode) == 'ADJOINT') then
      call outer_ad(1)
    else if (trim(mode) == 'FIXEDSRC') then
      call outer_fs(1)
    else
      call outer(1)
    end if
    allocate(npw(nnod))
    call PowDis(npw)
    allocate(dfout(nnod, ng, 6), dnout(nnod, ng, 6))
    do g = 1, ng
      do n = 1, nnod
        dfout(n, g, :) = nod(n, g)%df
        dnout(n, g, :) = nod(n, g)%dn
      end do
    end do
    write(u) Ke, f0, fs0, npw, dfout, dnout
    ! the reference's own CPU-time buckets of the solve (CMFD, nodal update;
    ! src/mod_cmfd.f90:442-489, :356-378), as ADPRES.f90 reports them
    write(*,'(A,2ES16.8)') ' REF_TIMES ', fdm_time, nod_time
    ! a fixed-source deck's external source (%ESRC, mod_io.f90:1375) follows
    if (trim(mode) == 'FIXEDSRC') write(u) exsrc

  case ('nodal')
    call read_state()
    call matrix_setup(1)
    do g = 1, ng
      do n = 1, nnod
        nod(n, g)%dn = dnin(n, g, :)
      end do
    end do
    ndmax = 0._dp
    if (kid == 1) then
      call nodal_update_pnm(1)
    else
      call nodal_update(1)
    end if
    allocate(dfout(nnod, ng, 6), dnout(nnod, ng, 6))
    do g = 1, ng
      do n = 1, nnod
        dfout(n, g, :) = nod(n, g)%df
        dnout(n, g, :) = nod(n, g)%dn
      end do
    end do
    write(u) dfout, dnout, ndmax, int(im, 4), int(jm, 4), int(km, 4)

  case ('bicg')
    call read_state()
    call matrix_setup(1)
    do g = 1, ng
      do n = 1, nnod
        nod(n, g)%dn = dnin(n, g, :)
      end do
    end do
    call matrix_setup(0)
    allocate(xx(nnod))
    do g = 1, ng
      xx = f0(:, g)
      call bicg(nin_p, g, bb(:, g), xx)
      f0(:, g) = xx
    end do
    write(u) f0

  case ('fwdadj')
    ! outer(0), then outer_ad(0) warm-started from the forward state (the
    ! sequence of rod_eject, src/mod_trans.f90:53-66, without KNE1)
    call outer(0)
    kes = Ke
    allocate(f0s(nnod, ng), fs0s(nnod))
    f0s = f0
    fs0s = fs0
    call outer_ad(0)
    allocate(dnout(nnod, ng, 6))
    do g = 1, ng
      do n = 1, nnod
        dnout(n, g, :) = nod(n, g)%dn
      end do
    end do
    write(u) kes, f0s, fs0s, Ke, f0, fs0, dnout, ndmax

  case ('xsupd')
    j = 0
    if (bbcon == 1 .or. bcbcs == 1) j = 1
    write(u) int(j, 4), int(bftem, 4), int(bmtem, 4), int(bcden, 4), int(bcrod, 4), int(nb, 4)
    write(u) rbcon, rftem, rmtem, rcden, coreh, pos0, ssize
    write(u) xsigtr, xsiga, xnuf, xsigf, xsigs
    if (j == 1) write(u) csigtr, csiga, cnuf, csigf, csigs
    if (bftem == 1) write(u) fsigtr, fsiga, fnuf, fsigf, fsigs
    if (bmtem == 1) write(u) msigtr, msiga, mnuf, msigf, msigs
    if (bcden == 1) write(u) lsigtr, lsiga, lnuf, lsigf, lsigs
    if (bcrod == 1) write(u) dsigtr, dsiga, dnuf, dsigf, dsigs, (int(fbmap(:, n), 4), n = 1, nyy)
    call write_xs_state()
    ! perturbed parameters: boron, node-wise temperatures and density, and
    ! every bank moved to a partial insertion
    bcon = bcon + 137._dp
    do n = 1, nnod
      if (bftem == 1) ftem(n) = ftem(n) + 40._dp * real(mod(n, 7) - 3, dp)
      if (bmtem == 1) mtem(n) = mtem(n) + 2._dp * real(mod(n, 5) - 2, dp)
      if (bcden == 1) cden(n) = cden(n) * (1._dp + 0.01_dp * real(mod(n, 3) - 1, dp))
    end do
    if (bcrod == 1) bpos = bpos * 0.5_dp + 17.25_dp
    call XS_updt(bcon, ftem, mtem, cden, bpos)
    call write_xs_state()

  case ('xtab')
    ! per material: branch dimensions, rod flag, branch points, then every
    ! branch point's cross sections (and the rodded set), s fastest
    write(u) int(bcrod, 4), int(nb, 4)
    write(u) (int(m(mm)%nd, 4), int(m(mm)%nb, 4), int(m(mm)%nf, 4), int(m(mm)%nm, 4), &
              int(m(mm)%trod, 4), mm = 1, nmat)
    do mm = 1, nmat
      write(u) m(mm)%pd, m(mm)%pb, m(mm)%pf, m(mm)%pm
      do v = 1, m(mm)%nm
        do q = 1, m(mm)%nf
          do t = 1, m(mm)%nb
            do s = 1, m(mm)%nd
              write(u) m(mm)%xsec(s, t, q, v)%sigtr, m(mm)%xsec(s, t, q, v)%siga, &
                       m(mm)%xsec(s, t, q, v)%nuf, m(mm)%xsec(s, t, q, v)%sigf, &
                       m(mm)%xsec(s, t, q, v)%sigs, m(mm)%xsec(s, t, q, v)%dc
            end do
          end do
        end do
      end do
      if (m(mm)%trod == 1) then
        do v = 1, m(mm)%nm
          do q = 1, m(mm)%nf
            do t = 1, m(mm)%nb
              do s = 1, m(mm)%nd
                write(u) m(mm)%rxsec(s, t, q, v)%sigtr, m(mm)%rxsec(s, t, q, v)%siga, &
                         m(mm)%rxsec(s, t, q, v)%nuf, m(mm)%rxsec(s, t, q, v)%sigf, &
                         m(mm)%rxsec(s, t, q, v)%sigs, m(mm)%rxsec(s, t, q, v)%dc
              end do
            end do
          end do
        end do
      end if
    end do
    write(u) coreh, pos0, ssize
    if (bcrod == 1) write(u) (int(fbmap(:, n), 4), n = 1, nyy)
    call write_xtab_state()
    ! perturbed parameters inside the branch ranges, every bank part-way in
    bcon = bcon + 37._dp
    do n = 1, nnod
      ftem(n) = ftem(n) * (1._dp + 0.05_dp * real(mod(n, 7) - 3, dp) / 3._dp)
      mtem(n) = mtem(n) + 1.5_dp * real(mod(n, 5) - 2, dp)
      cden(n) = cden(n) * (1._dp + 0.005_dp * real(mod(n, 3) - 1, dp))
    end do
    if (bcrod == 1) bpos = bpos * 0.5_dp + 17.25_dp
    call XStab_updt(bcon, ftem, mtem, cden, bpos)
    call write_xtab_state()

  case ('trstep')
    ! the steady state, then the set-up rod_eject and trans_calc make for a
    ! time step (src/mod_trans.f90:50-100, 400-416), with a rod-like
    ! removal increase in part of the core, then one outer_tr
    call outer(0)
    kes = Ke
    allocate(f0s(nnod, ng), fs0s(nnod), dns(nnod, ng, 6))
    f0s = f0
    fs0s = fs0
    do g = 1, ng
      do n = 1, nnod
        dns(n, g, :) = nod(n, g)%dn
      end do
    end do
    allocate(c0(nnod, nf), omeg(nnod, ng), L(nnod, ng))
    allocate(ft(nnod, ng), fst(nnod), dfis(nnod), sigrp(nnod, ng))
    tbeta = 0._dp
    do n = 1, nmat
      do j = 1, nf
        tbeta(n) = tbeta(n) + iBeta(j)
      end do
    end do
    do j = 1, nf
      do n = 1, nnod
        c0(n, j) = iBeta(j) / lamb(j) * fs0(n)
      end do
    end do
    omeg = 0._dp
    L = 0._dp
    ht = 0.25_dp
    ft = f0
    fst = fs0
    sigrp = sigr
    do g = 1, ng
      do n = 1, nnod
        if (2 * iz(n) > nzz .and. ix(n) <= 3) sigr(n, g) = sigr(n, g) * 1.002_dp
        sigr(n, g) = sigr(n, g) + 1._dp / (sth * velo(g) * ht)
      end do
    end do
    call outer_tr(ht, maxi)
    allocate(dfout(nnod, ng, 6), dnout(nnod, ng, 6))
    do g = 1, ng
      do n = 1, nnod
        dfout(n, g, :) = nod(n, g)%df
        dnout(n, g, :) = nod(n, g)%dn
      end do
    end do
    j = 0
    if (maxi) j = 1
    write(u) kes, f0s, fs0s, dns, ht, tbeta, sigr, dfis, exsrc, Ke, f0, fs0, dnout, dfout, int(j, 4)

  case default
    write(*,*) 'unknown action ', trim(action)
    stop 2
  end select

  close(u)

contains

  ! the parameters of an XS_updt call and the node cross sections it made
  subroutine write_xs_state()
    write(u) bcon
    if (bftem == 1) write(u) ftem
    if (bmtem == 1) write(u) mtem
    if (bcden == 1) write(u) cden
    if (bcrod == 1) write(u) bpos
    write(u) sigtr, siga, nuf, sigf, sigs, D, sigr
  end subroutine write_xs_state

  ! the parameters of an XStab_updt call and the node cross sections it made
  subroutine write_xtab_state()
    write(u) bcon, ftem, mtem, cden
    if (bcrod == 1) write(u) bpos
    write(u) sigtr, siga, nuf, sigf, sigs, D, sigr, dc
  end subroutine write_xtab_state

  subroutine read_state()
    ! STATEFILE: int32 nin, float64 Ke, f0(N,G), dn(N,G,6), b(N,G)
    integer :: us, itmp
    if (trim(stname) == '-') then
      write(*,*) 'action needs a STATEFILE'
      stop 2
    end if
    allocate(f0(nnod, ng), fs0(nnod), dnin(nnod, ng, 6), bb(nnod, ng))
    open(newunit=us, file=trim(stname), access='stream', form='unformatted', &
         status='old', action='read')
    read(us) itmp
    nin_p = itmp
    read(us) Ke, f0, dnin, bb
    close(us)
  end subroutine read_state

end program ref_dump
