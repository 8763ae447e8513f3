! Drop-in replacements for the reference's propagation and matching routines,
! with the reference's argument lists, calling libmcprop.so through the
! mcprop module.  A host program links these in place of its own
! MDPROP / LDPROP / MAPROP / AIPROP / SYMINV / YTOK and keeps everything
! else: the namelist readers, the BASE9 suites, POTENL (with VSTAR/VRTP),
! SCAPOT and PERTRB, DRCALC, the driver loops and the output.
!
! Hidden inputs are read where the reference reads them: NCONST, NRSQ,
! NVLBLK, CONLEN and VCONST from the host's module `potential`
! (source_code/potential_module.f:8-14), the VL indexing mode from
! COMMON /VLFLAG/ IVLFL and /VLSAVE/ IVLU (source_code/sumlam.f:22-23).
! Errors print the reference's message on unit 6 and STOP.

! P(lambda, R) for nr radii in HAMMAT's order (source_code/hammat.f:38-73):
! POTENL at R*RSCALE, the 1e60 clamp, the NCONST constants VCONST/EP2CM,
! SCAPOT, PERTRB at R*RSCALE, then EP2RU.  ptab(mxlam, nr) receives the
! radial terms; pconst(nconst) the constants, which must come out the same
! at every radius (a PERTRB shift of a constant that depends on R has no
! form in the library's per-instance constants and is refused).
subroutine mcpf_ptable(nr, r, mxlam, ep2ru, cm2ru, rscale, iprint, ptab, pconst)
  use, intrinsic :: iso_c_binding, only: c_double
  use potential, only: NCONST, NRSQ, VCONST
  implicit none
  integer, intent(in) :: nr, mxlam, iprint
  double precision, intent(in) :: r(nr), ep2ru, cm2ru, rscale
  real(c_double), intent(out) :: ptab(mxlam, nr)
  real(c_double), intent(out) :: pconst(max(NCONST, 1))
  double precision, allocatable :: p(:)
  integer :: k, i, idum1(1), idum2
  double precision :: ep2cm
  allocate (p(mxlam + NCONST + NRSQ + 1))
  ep2cm = ep2ru/cm2ru
  do k = 1, nr
    p = 0.0d0
    call POTENL(0, mxlam, idum1, r(k)*rscale, p, idum2, iprint)
    do i = 1, mxlam
      if (abs(p(i)) > 1.0d60) p(i) = sign(1.0d60, p(i))
    end do
    do i = 1, NCONST
      p(mxlam + i) = VCONST(i)/ep2cm
    end do
    call SCAPOT(p, mxlam)
    call PERTRB(r(k)*rscale, p, mxlam + NCONST + NRSQ, 0)
    do i = 1, mxlam + NCONST
      p(i) = ep2ru*p(i)
    end do
    ptab(:, k) = p(1:mxlam)
    if (k == 1) then
      pconst(1:NCONST) = p(mxlam + 1:mxlam + NCONST)
    else if (NCONST > 0) then
      if (any(pconst(1:NCONST) /= p(mxlam + 1:mxlam + NCONST))) then
        write (6, '(A)') ' *** ERROR IN mcprop: A CONSTANT TERM OF THE HAMILTONIAN VARIES WITH R ' // &
          '(PERTRB ON A CONSTANT WITH IPOW /= 0); NOT SUPPORTED'
        stop
      end if
    end if
  end do
  if (NCONST == 0) pconst(1) = 0.0d0
  deallocate (p)
end subroutine

! VL as the library takes it: NHAM rows (lambda fastest) per packed element
! of the upper triangle, column by column.  The reference's VL has NVLBLK >=
! NHAM rows; with IVLFL > 0 the entries map to P through IV
! (source_code/sumlam.f:68-86) and are summed per P index here.
subroutine mcpf_compact_vl(n, nham, vl, iv, vlc)
  use, intrinsic :: iso_c_binding, only: c_double
  use potential, only: NVLBLK
  implicit none
  integer, intent(in) :: n, nham
  double precision, intent(in) :: vl(*)
  integer, intent(in) :: iv(*)
  real(c_double), intent(out) :: vlc(nham, n*(n + 1)/2)
  integer :: ivlfl, ivlu, k, i, i1, nblk
  common /VLFLAG/ ivlfl
  common /VLSAVE/ ivlu
  if (ivlu /= 0) then
    write (6, '(A)') ' *** ERROR IN mcprop: VL HELD ON DISK (IVLU > 0) IS NOT SUPPORTED'
    stop
  end if
  nblk = max(NVLBLK, nham)
  vlc = 0.0d0
  do k = 1, n*(n + 1)/2
    i1 = (k - 1)*nblk
    if (ivlfl > 0) then
      do i = 1, nham
        if (vl(i1 + i) /= 0.0d0) vlc(iv(i1 + i), k) = vlc(iv(i1 + i), k) + vl(i1 + i)
      end do
    else
      vlc(:, k) = vl(i1 + 1:i1 + nham)
    end if
  end do
end subroutine

! One log-derivative segment (iprop 5 = LDPROP, 6 = MDPROP, 7 = MAPROP) on
! the grid R0 = RSTART, then the midpoint and end of each sector accumulated
! as R = R + H (source_code/mdprop.f:113,203), for one energy.
subroutine mcpf_ld_segment(iprop, n, mxlam, nham, y, vl, iv, eint, cent, rstart, nstep, dr, nodes, ered, &
                           ep2ru, cm2ru, rscale, iprint, istat)
  use, intrinsic :: iso_c_binding
  use mcprop
  use potential, only: NCONST, NRSQ, CONLEN
  implicit none
  integer, intent(in) :: iprop, n, mxlam, nham, nstep, iprint
  double precision, intent(inout) :: y(n*n)
  double precision, intent(in) :: vl(*), eint(n), cent(n), rstart, dr, ered, ep2ru, cm2ru, rscale
  integer, intent(in) :: iv(*)
  integer, intent(out) :: nodes, istat
  real(c_double), allocatable :: r(:), ptab(:), vlc(:), dummy(:)
  real(c_double), allocatable, target :: pconst(:)
  real(c_double) :: e1(1), h
  integer(c_int) :: ib, ig, nd(1), st(1)
  type(c_ptr) :: ctx, pc
  integer :: k, npts
  ctx = mcpf_context()
  npts = 2*nstep + 1
  allocate (r(npts), ptab(mxlam*npts), vlc(nham*(n*(n + 1)/2)), pconst(max(NCONST, 1)), dummy(0))
  h = 0.5d0*dr
  r(1) = rstart
  do k = 2, npts
    r(k) = r(k - 1) + h
  end do
  call mcpf_ptable(npts, r, mxlam, ep2ru, cm2ru, rscale, iprint, ptab, pconst)
  call mcpf_compact_vl(n, nham, vl, iv, vlc)
  ib = mcpf_block(n, mxlam, NCONST, NRSQ, vlc, eint, cent, CONLEN)
  ig = mcpf_grid(1, mxlam, npts, r, ptab, dr, dummy)
  pc = c_null_ptr
  if (NCONST > 0) pc = c_loc(pconst)
  e1(1) = ered
  call mcpf_check(mcp_prop_ld(ctx, ib, ig, int(iprop, c_int), 1_c_int, e1, pc, y, nd, st), 'PROPAGATOR')
  nodes = nd(1)
  istat = st(1)
end subroutine

subroutine MDPROP(N, MXLAM, NHAM, Y, VL, IV, EINT, CENT, RSTART, RSTOP, NSTEP, DR, NODES, IREC, WAVE, &
                  ERED, EP2RU, CM2RU, RSCALE, IPRINT)
  ! source_code/mdprop.f:1-4 (LDMD, IPROP = 6)
  implicit none
  integer :: N, MXLAM, NHAM, NSTEP, NODES, IREC, IPRINT, IV(*)
  double precision :: Y(N, N), VL(*), EINT(N), CENT(N), RSTART, RSTOP, DR, ERED, EP2RU, CM2RU, RSCALE
  logical :: WAVE
  integer :: istat
  if (WAVE) then
    write (6, '(A)') ' ***** MDPROP: WAVEFUNCTION RECORDS ARE NOT PRODUCED BY mcprop.  RUN HALTED.'
    stop
  end if
  call mcpf_ld_segment(6, N, MXLAM, NHAM, Y, VL, IV, EINT, CENT, RSTART, NSTEP, DR, NODES, ERED, EP2RU, CM2RU, &
                       RSCALE, IPRINT, istat)
  if (istat /= 0) then
    write (6, '(/A,I6,A)') ' ***** MDPROP: MATRIX INVERSION ERROR AT STEP K = ', istat, '.  RUN HALTED.'
    stop
  end if
end subroutine

subroutine LDPROP(N, MXLAM, NHAM, Z, VL, IV, EINT, CENT, RSTART, RSTOP, NSTEP, DR, NODES, ERED, EP2RU, CM2RU, &
                  RSCALE, IPRINT)
  ! source_code/ldprop.f:1-4 (LDJ, IPROP = 5)
  implicit none
  integer :: N, MXLAM, NHAM, NSTEP, NODES, IPRINT, IV(*)
  double precision :: Z(N, N), VL(*), EINT(N), CENT(N), RSTART, RSTOP, DR, ERED, EP2RU, CM2RU, RSCALE
  integer :: istat
  call mcpf_ld_segment(5, N, MXLAM, NHAM, Z, VL, IV, EINT, CENT, RSTART, NSTEP, DR, NODES, ERED, EP2RU, CM2RU, &
                       RSCALE, IPRINT, istat)
  if (istat /= 0) then
    write (6, '(/A)') ' *** ERROR IN SYMINV CALLED FROM LDPROP - TERMINATING'
    stop
  end if
end subroutine

subroutine MAPROP(N, MXLAM, NHAM, Y, VL, IV, EINT, CENT, RSTART, RSTOP, NSTEP, DR, NODES, ERED, EP2RU, CM2RU, &
                  RSCALE, IPRINT)
  ! source_code/maprop.f:1-4 (LDMA, IPROP = 7)
  implicit none
  integer :: N, MXLAM, NHAM, NSTEP, NODES, IPRINT, IV(*)
  double precision :: Y(N, N), VL(*), EINT(N), CENT(N), RSTART, RSTOP, DR, ERED, EP2RU, CM2RU, RSCALE
  integer :: istat
  call mcpf_ld_segment(7, N, MXLAM, NHAM, Y, VL, IV, EINT, CENT, RSTART, NSTEP, DR, NODES, ERED, EP2RU, CM2RU, &
                       RSCALE, IPRINT, istat)
  if (istat /= 0) then
    write (6, '(/A,I6,A)') ' ***** MATRIX INVERSION ERROR IN MAPROP AT STEP  K = ', istat, '  RUN HALTED.'
    stop
  end if
end subroutine

subroutine AIPROP(NCH, MXLAM, NHAM, Z, VL, IV, EINT, CENT, RSTART, RSTOP, NSTEP, DRNOW, POWR, TOLHI, NODES, &
                  ERED, EP2RU, CM2RU, RSCALE, IPRINT, IREC, WAVE)
  ! source_code/aiprop.f:1-4 (AIRY, IPROP = 9).  TOLHI = 0: the steps
  ! follow DRCALC's R**POWR series from DRNOW, NSTEP of them (aiprop.f:
  ! 288-327), the potential at each step's two Gauss points RNOW -+
  ! DRNOW/(2 sqrt 3) (source_code/potent.f:66-67).  0 < TOLHI < 1: the
  ! steps are chosen as the propagation goes, which needs the &POTL terms
  ! registered with mcpf_set_potl.  NSTEP returns the steps taken.
  use, intrinsic :: iso_c_binding
  use mcprop
  use potential, only: NCONST, NRSQ, CONLEN
  implicit none
  integer :: NCH, MXLAM, NHAM, NSTEP, NODES, IPRINT, IREC, IV(*)
  double precision :: Z(NCH, NCH), VL(*), EINT(NCH), CENT(NCH), RSTART, RSTOP, DRNOW, POWR, TOLHI, ERED, EP2RU, &
                      CM2RU, RSCALE
  logical :: WAVE
  double precision, external :: DRCALC
  real(c_double), allocatable :: r(:), ptab(:), vlc(:), widths(:), rnows(:), steps(:)
  real(c_double) :: pconst(1), e1(1), dr, rlast, rnext, rnow
  integer(c_int) :: ib, ig, nd(1), st(1), ns(1), nrec
  type(c_ptr) :: ctx
  integer :: k, kstep, npts, mxstep
  logical :: last
  double precision, parameter :: SQ3 = 1.732050807d0
  if (WAVE) then
    write (6, '(A)') ' *** AIPROP: WAVEFUNCTION RECORDS ARE NOT PRODUCED BY mcprop.  RUN HALTED.'
    stop
  end if
  if (NCONST /= 0 .or. NRSQ /= 0) then
    write (6, '(A)') ' *** AIPROP: mcprop TAKES BLOCKS WITH NCONST = NRSQ = 0 ONLY.  RUN HALTED.'
    stop
  end if
  ctx = mcpf_context()
  allocate (vlc(NHAM*(NCH*(NCH + 1)/2)))
  call mcpf_compact_vl(NCH, NHAM, VL, IV, vlc)
  ib = mcpf_block(NCH, MXLAM, NCONST, NRSQ, vlc, EINT, CENT, CONLEN)
  e1(1) = ERED
  if (TOLHI <= 0.0d0) then
    ! the step widths and midpoints in AIPROP's order
    allocate (widths(NSTEP + 1), rnows(NSTEP + 1))
    dr = DRNOW
    rlast = RSTART
    rnow = rlast + 0.5d0*dr
    rnext = rlast + dr
    last = .false.
    kstep = 0
    do
      kstep = kstep + 1
      rnows(kstep) = rnow
      widths(kstep) = dr
      if (last) exit
      dr = DRCALC(RSTART, RSTOP, kstep, NSTEP, POWR)
      rlast = rnext
      rnext = rnext + dr
      last = (NSTEP - kstep) <= 1
      if (last) then
        rnext = RSTOP
        dr = rnext - rlast
      end if
      rnow = rlast + 0.5d0*dr
    end do
    npts = 2*kstep + 1
    allocate (r(npts), ptab(MXLAM*npts))
    r(1) = RSTART
    do k = 1, kstep
      r(2*k) = rnows(k) - 0.5d0*widths(k)/SQ3
      r(2*k + 1) = rnows(k) + 0.5d0*widths(k)/SQ3
    end do
    call mcpf_ptable(npts, r, MXLAM, EP2RU, CM2RU, RSCALE, IPRINT, ptab, pconst)
    ig = mcpf_grid(2, MXLAM, npts, r, ptab, 0.0d0, widths(1:kstep))
    call mcpf_check(mcp_prop_airy(ctx, ib, ig, 1_c_int, e1, Z, nd, st), 'AIPROP')
    NSTEP = kstep
  else
    if (mcpf_potl_nterm < 0) then
      write (6, '(A)') ' *** AIPROP: TOLHI > 0 NEEDS THE &POTL TERMS (mcpf_set_potl).  RUN HALTED.'
      stop
    end if
    call mcpf_check(mcp_block_set_potl(ctx, ib, int(mcpf_potl_nterm, c_int), mcpf_potl_lam, mcpf_potl_npower, &
                                       mcpf_potl_a, mcpf_potl_e, RSCALE, EP2RU*mcpf_scalam), 'AIPROP')
    mxstep = 100000
    allocate (steps(2*mxstep))
    nrec = 0
    call mcpf_check(mcp_prop_airy_tolhi(ctx, ib, 1_c_int, e1, Z, RSTART, RSTOP, DRNOW, POWR, TOLHI, 0_c_int, &
                                        int(mxstep, c_int), steps, nrec, nd, ns, st), 'AIPROP')
    NSTEP = ns(1)
    if (st(1) == -1) then
      write (6, '(A)') ' *** AIRY PROPAGATION NOT FINISHED.'
      st(1) = 0
    end if
  end if
  if (st(1) /= 0) then
    write (6, '(A)') '  *** INSTABILITY IN SYMINV IN AIPROP.'
    stop
  end if
  NODES = nd(1)
end subroutine

subroutine VVPROP(N, MXLAM, NHAM, RMAT, VL, IVL, EINT, CENT, RSTART, RSTOP, NSTEP, DRNOW, DRMAX, TLDIAG, TOFF, &
                  ERED, EP2RU, CM2RU, RSCALE, IPRINT)
  ! source_code/vvprop.f:1-4 (VIVS, IPROP = 4): RMAT, the R-matrix, from
  ! RSTART to RSTOP under the /LDVVCM/ controls (vvprop.f:57-82), the
  ! potential from the &POTL terms registered with mcpf_set_potl (the device
  ! evaluates them, and DERMAT's derivatives, where the step control lands).
  ! /PRPSCR/'s IWRITE keeps this call's steps, IREAD replays the kept steps
  ! (the scratch file ISCRU of the energies after the first, vvprop.f:
  ! 150-160, 350-352, 590-593).  NSTEP returns the steps taken.
  use, intrinsic :: iso_c_binding
  use mcprop
  use potential, only: NCONST, NRSQ, CONLEN
  implicit none
  integer :: N, MXLAM, NHAM, NSTEP, IPRINT, IVL(*)
  double precision :: RMAT(N, N), VL(*), EINT(N), CENT(N), RSTART, RSTOP, DRNOW, DRMAX, TLDIAG, TOFF, ERED, &
                      EP2RU, CM2RU, RSCALE
  double precision :: XSQMAX, ALPHA1, ALPHA2, ESHIFT
  integer :: IALPHA, ISCRU, ISCRUR
  logical :: IALFP, IV, IVP, IVPP, ISHIFT, IDIAG, IPERT, ISYM, IREAD, IWRITE, IREADR, IWRITR
  common /LDVVCM/ XSQMAX, ALPHA1, ALPHA2, IALPHA, IALFP, IV, IVP, IVPP, ISHIFT, IDIAG, IPERT, ISYM
  common /PRPSCR/ ESHIFT, ISCRU, ISCRUR, IREAD, IWRITE, IREADR, IWRITR
  real(c_double), allocatable :: vlc(:), steps(:)
  real(c_double) :: e1(1)
  integer(c_int) :: ib, st(1), ns(1), replay, mxstep
  type(c_ptr) :: ctx
  if (NCONST /= 0 .or. NRSQ /= 0) then
    write (6, '(A)') ' *** VVPROP: mcprop TAKES BLOCKS WITH NCONST = NRSQ = 0 ONLY.  RUN HALTED.'
    stop
  end if
  if (mcpf_potl_nterm < 0) then
    write (6, '(A)') ' *** VVPROP: VIVS NEEDS THE &POTL TERMS (mcpf_set_potl).  RUN HALTED.'
    stop
  end if
  if (IREAD .and. mcpf_vivs_nrec <= 0) then
    write (6, '(A)') ' *** VVPROP: IREAD WITHOUT STEPS KEPT BY AN EARLIER IWRITE CALL.  RUN HALTED.'
    stop
  end if
  ctx = mcpf_context()
  allocate (vlc(NHAM*(N*(N + 1)/2)))
  call mcpf_compact_vl(N, NHAM, VL, IVL, vlc)
  ib = mcpf_block(N, MXLAM, NCONST, NRSQ, vlc, EINT, CENT, CONLEN)
  call mcpf_check(mcp_block_set_potl(ctx, ib, int(mcpf_potl_nterm, c_int), mcpf_potl_lam, mcpf_potl_npower, &
                                     mcpf_potl_a, mcpf_potl_e, RSCALE, EP2RU*mcpf_scalam), 'VVPROP')
  call mcpf_check(mcp_set_vivs_options(ctx, l2i(IV), l2i(IVP), l2i(IVPP), l2i(ISHIFT), l2i(IDIAG), l2i(IPERT), &
                                       l2i(ISYM), int(IALPHA, c_int), l2i(IALFP), ALPHA1, ALPHA2, XSQMAX), 'VVPROP')
  e1(1) = ERED
  if (IREAD) then
    mxstep = int(mcpf_vivs_nrec, c_int)
    replay = 1
    allocate (steps(4*mxstep))
    steps = mcpf_vivs_rec(1:4*mxstep)
  else
    mxstep = 200000
    replay = 0
    allocate (steps(4*mxstep))
  end if
  call mcpf_check(mcp_prop_vivs(ctx, ib, 1_c_int, e1, RMAT, RSTART, RSTOP, DRNOW, DRMAX, TLDIAG, TOFF, replay, &
                                mxstep, steps, ns, st), 'VVPROP')
  if (st(1) == 1) then
    write (6, '(A)') ' *** VVPROP: DELRD STEP SIZE TOO SMALL.  RUN HALTED.'
    stop
  else if (st(1) /= 0) then
    write (6, '(A,I3,A)') ' *** VVPROP: PROPAGATION FAILED (STATUS', st(1), ').  RUN HALTED.'
    stop
  end if
  NSTEP = ns(1)
  if (IWRITE .and. .not. IREAD) then
    mcpf_vivs_rec = steps(1:4*NSTEP)
    mcpf_vivs_nrec = NSTEP
  end if
contains
  integer(c_int) function l2i(b)
    logical, intent(in) :: b
    l2i = 0
    if (b) l2i = 1
  end function
end subroutine

subroutine BDCTRL(N, MXLAM, NHAM, PSIMCH, VL, IV, EINT, CENT, NODEC, ERED, EP2RU, CM2RU, RSCALE, EIGMIN, WAVE, &
                  IMIN, IPRINT)
  ! source_code/bdctrl.f:1-4: BOUND's matching at one energy for the
  ! two-segment layout with RMID = RMATCH (source_code/bd.prpini.f:24-62,
  ! bdpset.f:26-42): the outward segment RMIN -> RMATCH with DRS and the
  ! inward one RMAX -> RMATCH with DRL, each at DRSET's constant step
  ! (STEP = -10, drset.f:62-64), IPROPS outward and IPROPL inward (5, 6 or
  ! 7), read from /RADIAL/; the matching matrix's eigenvalues, NODEC (both
  ! propagations' node counts plus the negative eigenvalues), EIGMIN, IMIN
  ! and PSIMCH from mcp_bound_match.  The wavefunction (WAVE) is
  ! mcp_bound_wave's, not this drop-in's.
  use, intrinsic :: iso_c_binding
  use mcprop
  use potential, only: NCONST, NRSQ, CONLEN
  implicit none
  integer :: N, MXLAM, NHAM, NODEC, IMIN, IPRINT, IV(*)
  double precision, target :: PSIMCH(N)
  double precision :: VL(*), EINT(N), CENT(N), ERED, EP2RU, CM2RU, RSCALE, EIGMIN
  logical :: WAVE
  double precision :: RMNINT, RMXINT, RMID, RMATCH, DRS, DRL, STEPS, STEPL, POWRS, POWRL, TOLHIS, TOLHIL, &
                      CAYS, CAYL, UNSET
  integer :: IPROPS, IPROPL, NSEG
  common /RADIAL/ RMNINT, RMXINT, RMID, RMATCH, DRS, DRL, STEPS, STEPL, POWRS, POWRL, TOLHIS, TOLHIL, CAYS, CAYL, &
                  UNSET, IPROPS, IPROPL, NSEG
  real(c_double), allocatable :: vlc(:), ro(:), ri(:), pto(:), pti(:), dummy(:)
  real(c_double), allocatable, target :: pconst(:), evals(:)
  real(c_double) :: e1(1), em(1), dro, dri
  integer(c_int) :: ib, igo, igi, nd(1), im(1), st(1)
  type(c_ptr) :: ctx, pc
  integer :: nout, nin, k
  double precision :: drl_use
  if (WAVE) then
    write (6, '(A)') ' *** BDCTRL: THE WAVEFUNCTION IS mcp_bound_wave''S, NOT THIS DROP-IN''S.  RUN HALTED.'
    stop
  end if
  if (IPROPS < 5 .or. IPROPS > 7 .or. IPROPL < 5 .or. IPROPL > 7) then
    write (6, '(A,2I3,A)') ' *** BDCTRL: IPROPS, IPROPL =', IPROPS, IPROPL, ' (5, 6 OR 7 HERE).  RUN HALTED.'
    stop
  end if
  drl_use = DRL
  if (DRL == UNSET) drl_use = DRS
  ! DRSET's constant step: NSTEP = max(nint(range / DR), 1), DR = range / NSTEP
  nout = max(int(floor(abs(RMATCH - RMNINT)/DRS + 0.5d0)), 1)
  dro = abs(RMATCH - RMNINT)/dble(nout)
  nin = max(int(floor(abs(RMXINT - RMATCH)/drl_use + 0.5d0)), 1)
  dri = -abs(RMXINT - RMATCH)/dble(nin)
  allocate (ro(2*nout + 1), ri(2*nin + 1), pto(MXLAM*(2*nout + 1)), pti(MXLAM*(2*nin + 1)), dummy(0), &
            vlc(NHAM*(N*(N + 1)/2)), pconst(max(NCONST, 1)), evals(N))
  ro(1) = RMNINT
  do k = 2, 2*nout + 1
    ro(k) = ro(k - 1) + 0.5d0*dro
  end do
  ri(1) = RMXINT
  do k = 2, 2*nin + 1
    ri(k) = ri(k - 1) + 0.5d0*dri
  end do
  call mcpf_ptable(2*nout + 1, ro, MXLAM, EP2RU, CM2RU, RSCALE, IPRINT, pto, pconst)
  call mcpf_ptable(2*nin + 1, ri, MXLAM, EP2RU, CM2RU, RSCALE, IPRINT, pti, pconst)
  ctx = mcpf_context()
  call mcpf_compact_vl(N, NHAM, VL, IV, vlc)
  ib = mcpf_block(N, MXLAM, NCONST, NRSQ, vlc, EINT, CENT, CONLEN)
  igo = mcpf_grid(1, MXLAM, 2*nout + 1, ro, pto, dro, dummy)
  igi = mcpf_grid(1, MXLAM, 2*nin + 1, ri, pti, dri, dummy)
  pc = c_null_ptr
  if (NCONST > 0) pc = c_loc(pconst)
  e1(1) = ERED
  call mcpf_check(mcp_bound_match(ctx, ib, igo, igi, int(IPROPS, c_int), int(IPROPL, c_int), 1_c_int, e1, pc, &
                                  nd, em, im, c_loc(PSIMCH), c_loc(evals), st), 'BDCTRL')
  if (st(1) /= 0) then
    write (6, '(A,I9,A)') ' *** BDCTRL: PROPAGATION FAILED (STATUS', st(1), ').  RUN HALTED.'
    stop
  end if
  NODEC = nd(1)
  EIGMIN = em(1)
  IMIN = im(1) + 1
end subroutine

subroutine SYMINV(A, IA, N, IFAIL)
  ! source_code/syminv.f:1 (with SYMINL for N > 30): the lower triangle of
  ! A(1:N,1:N) in, the full symmetric inverse out; IFAIL = the number of
  ! negative eigenvalues, N+1 for an exactly zero pivot (N <= 30), -INFO for
  ! an exactly singular column (N > 30).
  use, intrinsic :: iso_c_binding
  use mcprop
  implicit none
  integer :: IA, N, IFAIL
  double precision :: A(IA, N)
  real(c_double), allocatable :: b(:, :)
  integer(c_int) :: kount(1)
  if (N <= 0) then
    IFAIL = 0
    return
  end if
  allocate (b(N, N))
  b = A(1:N, 1:N)
  call mcpf_check(mcp_syminv_batched(mcpf_context(), int(N, c_int), 1_c_int, b, b, kount), 'SYMINV')
  A(1:N, 1:N) = b
  IFAIL = kount(1)
end subroutine

subroutine YTOK(NB, WVEC, L, N, NOPEN, Y, KMTRIX, RVAL, CENT)
  ! source_code/ytok.f:1 (+ YTOKG): K (NOPEN x NOPEN, leading dimension
  ! NOPEN) from the lower triangle of Y at RVAL.  KTOS stays the host's
  ! (an NOPEN x NOPEN computation).
  use, intrinsic :: iso_c_binding
  use mcprop
  implicit none
  integer :: N, NOPEN, NB(N), L(N)
  double precision :: WVEC(N), Y(N, N), KMTRIX(*), RVAL, CENT(N)
  real(c_double), allocatable :: yy(:, :), kk(:), sr(:), si(:)
  integer(c_int), allocatable :: nb0(:)
  integer(c_int) :: nop(1), st(1)
  integer :: i, j
  if (NOPEN == 0) return
  allocate (yy(N, N), kk(N*N), sr(N*N), si(N*N), nb0(N))
  do j = 1, N
    do i = j, N
      yy(i, j) = Y(i, j)
      yy(j, i) = Y(i, j)
    end do
  end do
  nb0 = int(NB - 1, c_int)
  nop(1) = int(NOPEN, c_int)
  call mcpf_check(mcp_ytok_ktos(mcpf_context(), int(N, c_int), 1_c_int, yy, WVEC, nop, nb0, CENT, RVAL, kk, sr, &
                                si, st), 'YTOK')
  if (st(1) /= 0) then
    write (6, '(A,I6)') ' *** ERROR IN YTOK: MATCHING FAILED, CODE ', st(1)
    stop
  end if
  KMTRIX(1:NOPEN*NOPEN) = kk(1:NOPEN*NOPEN)
end subroutine
