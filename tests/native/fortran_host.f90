! Test host program for the Fortran drop-ins (molscat_amd/fortran): it plays
! the part of the reference's host program around the propagators, with
! its own module `potential`, POTENL (a sum of power / exponential terms per
! radial coefficient, the &POTL form), SCAPOT, PERTRB and DRCALC, and drives
! MDPROP / LDPROP / MAPROP / AIPROP / YTOK / SYMINV / VVPROP / BDCTRL through their
! reference argument lists.
!
!   fortran_host cpu                 checks that need no GPU
!   fortran_host run <in> <out>      one deck through the drop-ins
module potential
  implicit none
  integer :: NVLBLK = 0, NCONST = 0, NRSQ = 0
  double precision :: CONLEN = 0.0d0, VCONST(16) = 0.0d0
end module potential

module host_potl
  implicit none
  integer :: nterm = 0
  integer, allocatable :: tlam(:), tpow(:)
  double precision, allocatable :: ta(:), te(:)
  double precision :: scalam = 1.0d0
end module host_potl

subroutine POTENL(icntrl, mxlmb, lam, r, p, itype, iprint)
  use host_potl
  implicit none
  integer :: icntrl, mxlmb, lam(*), itype, iprint
  double precision :: r, p(*), x
  integer :: t
  p(1:mxlmb) = 0.0d0
  do t = 1, nterm
    if (tpow(t) /= 0) then
      x = r
      if (tpow(t) < 0 .and. abs(x) < 1.0d-8) x = 1.0d-8
      p(tlam(t)) = p(tlam(t)) + ta(t)*x**tpow(t)
    else
      p(tlam(t)) = p(tlam(t)) + ta(t)*exp(te(t)*r)
    end if
  end do
end subroutine

subroutine SCAPOT(p, mxlam)
  use host_potl, only: scalam
  implicit none
  integer :: mxlam
  double precision :: p(mxlam)
  p = p*scalam
end subroutine

subroutine PERTRB(r, p, nham, ideriv)
  implicit none
  integer :: nham, ideriv
  double precision :: r, p(nham)
  integer :: i, ipow
  double precision :: deltan
  common /EXPVAL/ i, ipow, deltan
  if (i > 0 .and. i <= nham) p(i) = p(i) + deltan/r**ipow
end subroutine

! Width of step kstep+1 of nstep steps from rstart to rstop growing as
! R**pow (constant in log R for pow = 1).
double precision function DRCALC(rstart, rstop, kstep, nstep, pow)
  implicit none
  double precision :: rstart, rstop, pow
  integer :: kstep, nstep
  double precision :: q, z0, dz
  q = 1.0d0 - pow
  if (q /= 0.0d0) then
    z0 = rstart**q
    dz = (rstop**q - z0)/dble(nstep)
    DRCALC = (z0 + dble(kstep + 1)*dz)**(1.0d0/q) - (z0 + dble(kstep)*dz)**(1.0d0/q)
  else
    dz = rstart*((rstop/rstart)**(1.0d0/dble(nstep)) - 1.0d0)
    DRCALC = dz*(1.0d0 + dz/rstart)**kstep
  end if
end function

program fortran_host
  use, intrinsic :: iso_c_binding
  use mcprop
  implicit none
  character(len=512) :: mode
  call get_command_argument(1, mode)
  if (trim(mode) == 'cpu') then
    call cpu_checks()
  else if (trim(mode) == 'run') then
    call run_deck()
  else
    write (6, '(A)') 'usage: fortran_host cpu | run <in> <out>'
    stop 2
  end if

contains

  subroutine cpu_checks()
    integer(c_int) :: ns, rc, id, kount(1), nd(1), st(1)
    real(c_double) :: a(8), b(8), m(4), e1(1)
    integer :: k
    rc = mcp_mg_scheme(4_c_int, ns, a, b)
    write (6, '(A,2I4,12ES25.16)') 'MG4', rc, ns, (a(k), k=1, ns), (b(k), k=1, ns)
    rc = mcp_mg_scheme(5_c_int, ns, a, b)
    write (6, '(A,2I4,12ES25.16)') 'MG5', rc, ns, (a(k), k=1, ns), (b(k), k=1, ns)
    rc = mcp_mg_scheme(3_c_int, ns, a, b)
    write (6, '(A,I4)') 'MG3', rc
    m = [2.0d0, 0.0d0, 0.0d0, 2.0d0]
    e1(1) = 0.0d0
    rc = mcp_syminv_batched(c_null_ptr, 2_c_int, 1_c_int, m, m, kount)
    write (6, '(A,I4)') 'NULLCTX syminv', rc
    rc = mcp_block_create(c_null_ptr, 2_c_int, 1_c_int, 0_c_int, 0_c_int, m, m, m, 0.0d0, id)
    write (6, '(A,I4)') 'NULLCTX block_create', rc
    rc = mcp_prop_ld(c_null_ptr, 0_c_int, 0_c_int, 6_c_int, 1_c_int, e1, c_null_ptr, m, nd, st)
    write (6, '(A,I4)') 'NULLCTX prop_ld', rc
    rc = mcp_bound_match(c_null_ptr, 0_c_int, 0_c_int, 0_c_int, 6_c_int, 6_c_int, 1_c_int, e1, c_null_ptr, nd, &
                         m, nd, c_null_ptr, c_null_ptr, st)
    write (6, '(A,I4)') 'NULLCTX bound_match', rc
    write (6, '(A)') 'CPU_CHECKS_DONE'
  end subroutine

  subroutine run_deck()
    use potential
    use host_potl
    character(len=512) :: fin, fout
    integer :: n, mxlam, nham, nstep, nstepa, nopen, nodes, nodes2, nodes_ld, nodes_ma, nodes_a, ifail, i, j, t
    integer :: nb_c, ng_c, nup_b, nup_g, nup_b1, nup_g1, nsteps_used, nstep2
    ! the /LDVVCM/ and /PRPSCR/ common blocks (local names prefixed with c:
    ! IV is also the host's coupling index array here)
    double precision :: cxsqmax, calpha1, calpha2, ceshift
    integer :: cialpha, ciscru, ciscrur
    logical :: cialfp, civ, civp, civpp, cishift, cidiag, cipert, cisym, ciread, ciwrite, cireadr, ciwritr
    common /LDVVCM/ cxsqmax, calpha1, calpha2, cialpha, cialfp, civ, civp, civpp, cishift, cidiag, cipert, cisym
    common /PRPSCR/ ceshift, ciscru, ciscrur, ciread, ciwrite, cireadr, ciwritr
    double precision :: crmnint, crmxint, crmid, crmatch, cdrs, cdrl, csteps, cstepl, cpowrs, cpowrl, ctolhis, &
                        ctolhil, ccays, ccayl, cunset
    integer :: ciprops, ciprorl, cnseg, nodec_b, imin_b
    double precision :: eigmin_b
    double precision, allocatable :: psi_b(:)
    common /RADIAL/ crmnint, crmxint, crmid, crmatch, cdrs, cdrl, csteps, cstepl, cpowrs, cpowrl, ctolhis, ctolhil, &
                    ccays, ccayl, cunset, ciprops, ciprorl, cnseg
    double precision :: ep2ru, cm2ru, rscale, ered, rstart, rmid, dr, drnow, powr, rmax
    double precision, allocatable :: vl(:), eint(:), cent(:), y0(:, :), y(:, :), z(:, :), y2(:, :), wvec(:), &
                                     kmat(:), a(:, :)
    integer, allocatable :: nb(:), lv(:), iv(:)
    integer :: u
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    open (newunit=u, file=trim(fin), status='old', action='read')
    read (u, *) n, mxlam, nham
    NVLBLK = nham
    allocate (vl(nham*n*(n + 1)/2), eint(n), cent(n), y0(n, n), y(n, n), z(n, n), y2(n, n), wvec(n), &
              kmat(n*n), nb(n), lv(n), iv(1), a(n, n), psi_b(n))
    read (u, *) vl
    read (u, *) eint
    read (u, *) cent
    read (u, *) nterm
    allocate (tlam(nterm), tpow(nterm), ta(nterm), te(nterm))
    do t = 1, nterm
      read (u, *) tlam(t), tpow(t), ta(t), te(t)
    end do
    read (u, *) scalam
    read (u, *) ep2ru, cm2ru, rscale
    read (u, *) ered
    read (u, *) rstart, rmid, nstep, dr
    read (u, *) nstepa, drnow, powr, rmax
    read (u, *) y0
    read (u, *) nopen
    read (u, *) nb
    read (u, *) wvec
    close (u)
    lv = 0
    iv = 0

    open (newunit=u, file=trim(fout), status='replace', action='write')
    ! MDPROP twice at the same energy: the second call reuses the uploads
    y = y0
    call MDPROP(n, mxlam, nham, y, vl, iv, eint, cent, rstart, rmid, nstep, dr, nodes, 0, .false., ered, ep2ru, &
                cm2ru, rscale, 0)
    call mcpf_cache_stats(nb_c, ng_c, nup_b1, nup_g1)
    y2 = y0
    call MDPROP(n, mxlam, nham, y2, vl, iv, eint, cent, rstart, rmid, nstep, dr, nodes2, 0, .false., ered, ep2ru, &
                cm2ru, rscale, 0)
    call mcpf_cache_stats(nb_c, ng_c, nup_b, nup_g)
    write (u, '(A,4I8)') 'CACHE', nup_b1, nup_g1, nup_b, nup_g
    write (u, '(A,I10,I10,ES25.16)') 'MDPROP', nodes, nodes2, maxval(abs(y - y2))
    write (u, '(ES25.16)') y
    z = y0
    call LDPROP(n, mxlam, nham, z, vl, iv, eint, cent, rstart, rmid, nstep, dr, nodes_ld, ered, ep2ru, cm2ru, &
                rscale, 0)
    write (u, '(A,I10)') 'LDPROP', nodes_ld
    write (u, '(ES25.16)') z
    z = y0
    call MAPROP(n, mxlam, nham, z, vl, iv, eint, cent, rstart, rmid, nstep, dr, nodes_ma, ered, ep2ru, cm2ru, &
                rscale, 0)
    write (u, '(A,I10)') 'MAPROP', nodes_ma
    write (u, '(ES25.16)') z
    ! AIRY (TOLHI = 0) from the MDPROP result to RMAX
    z = y
    nsteps_used = nstepa
    call AIPROP(n, mxlam, nham, z, vl, iv, eint, cent, rmid, rmax, nsteps_used, drnow, powr, 0.0d0, nodes_a, &
                ered, ep2ru, cm2ru, rscale, 0, 0, .false.)
    write (u, '(A,I10,I10)') 'AIPROP', nodes_a, nsteps_used
    write (u, '(ES25.16)') z
    ! YTOK at RMAX
    call YTOK(nb, wvec, lv, n, nopen, z, kmat, rmax, cent)
    write (u, '(A,I6)') 'YTOK', nopen
    write (u, '(ES25.16)') kmat(1:nopen*nopen)
    ! SYMINV of the MDPROP result (lower triangle in, full inverse out)
    a = y
    do j = 1, n
      do i = 1, j - 1
        a(i, j) = 1.0d300
      end do
    end do
    call SYMINV(a, n, n, ifail)
    write (u, '(A,I6)') 'SYMINV', ifail
    write (u, '(ES25.16)') a
    ! VVPROP from RMID to RMAX on R = Y^-1 under IDIAG with DELRD steps
    ! (IALPHA = 0): once choosing and keeping its steps (IWRITE), once
    ! replaying them (IREAD)
    call mcpf_set_potl(nterm, tlam, tpow, ta, te)
    cxsqmax = 1.0d4
    calpha1 = 1.0d0
    calpha2 = 1.5d0
    cialpha = 0
    cialfp = .false.
    civ = .true.
    civp = .true.
    civpp = .true.
    cishift = .true.
    cidiag = .true.
    cipert = .true.
    cisym = .true.
    ciread = .false.
    ciwrite = .true.
    y = a
    call VVPROP(n, mxlam, nham, y, vl, iv, eint, cent, rmid, rmax, nsteps_used, 0.002d0, 5.0d0, 0.064d0, 0.064d0, &
                ered, ep2ru, cm2ru, rscale, 0)
    ciread = .true.
    ciwrite = .false.
    y2 = a
    call VVPROP(n, mxlam, nham, y2, vl, iv, eint, cent, rmid, rmax, nstep2, 0.002d0, 5.0d0, 0.064d0, 0.064d0, &
                ered, ep2ru, cm2ru, rscale, 0)
    write (u, '(A,I10,I10,ES25.16)') 'VVPROP', nsteps_used, nstep2, maxval(abs(y - y2))
    write (u, '(ES25.16)') y
    ! BDCTRL: MDPROP outward RSTART -> RMID and inward RMAX -> RMID at ERED
    crmnint = rstart
    crmxint = rmax
    crmid = rmid
    crmatch = rmid
    cdrs = dr
    cdrl = 2.0d0*dr
    cunset = -1.0d30
    ciprops = 6
    ciprorl = 6
    cnseg = 2
    call BDCTRL(n, mxlam, nham, psi_b, vl, iv, eint, cent, nodec_b, ered, ep2ru, cm2ru, rscale, eigmin_b, &
                .false., imin_b, 0)
    write (u, '(A,I10,I6,ES25.16)') 'BDCTRL', nodec_b, imin_b, eigmin_b
    write (u, '(ES25.16)') psi_b
    close (u)
    call mcpf_release()
    write (6, '(A)') 'RUN_DONE'
  end subroutine

end program fortran_host
