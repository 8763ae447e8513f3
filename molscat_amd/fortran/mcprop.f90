! mcprop: ISO_C_BINDING interfaces to libmcprop.so (include/mcprop.h) for a
! Fortran host program, plus the state the drop-in propagators keep between
! calls: one context, and caches of the device copies of blocks (VL, EINT,
! CENT) and radial grids (radii and P tables).
!
! The reference keeps its propagation state in COMMON blocks and allocates
! scratch per call (source_code/mdprop.f:56-57,76,263-265).  Here the
! device objects live in the library's context; a block or grid is uploaded
! once and reused while later calls pass the same data (the energies of one
! (JTOT, IB) block, or every block on one grid), and the least recently used
! entry is destroyed when a new one needs its slot, so a long run holds at
! most MCPF_NCACHE blocks and MCPF_NCACHE grids on the device.
module mcprop
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: MCP_OK = 0
  integer(c_int), parameter :: MCP_ERR_ARG = -1
  integer(c_int), parameter :: MCP_ERR_HIP = -2
  integer(c_int), parameter :: MCP_ERR_UNSUPPORTED = -3
  integer(c_int), parameter :: MCP_ERR_NOMEM = -4
  integer(c_int), parameter :: MCP_PROP_ODPROP = 106

  interface
    integer(c_int) function mcp_ctx_create(device, ctx) bind(C, name="mcp_ctx_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
    end function
    subroutine mcp_ctx_destroy(ctx) bind(C, name="mcp_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function mcp_last_error(ctx) bind(C, name="mcp_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcp_ctx_info(ctx, max_n_md, max_n_ld, max_n_eig, lds_bytes, num_cu) &
        bind(C, name="mcp_ctx_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: max_n_md, max_n_ld, max_n_eig, lds_bytes, num_cu
    end function
    integer(c_int) function mcp_syminv_batched(ctx, n, nmat, a, ainv, kount) bind(C, name="mcp_syminv_batched")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, nmat
      real(c_double) :: a(*), ainv(*)
      integer(c_int) :: kount(*)
    end function
    integer(c_int) function mcp_eigh_batched(ctx, n, nmat, a, evals, evecs) bind(C, name="mcp_eigh_batched")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, nmat
      real(c_double) :: a(*), evals(*), evecs(*)
    end function
    integer(c_int) function mcp_block_create(ctx, n, mxlam, nconst, nrsq, vl, eint, cent, conlen, id) &
        bind(C, name="mcp_block_create")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, mxlam, nconst, nrsq
      real(c_double) :: vl(*), eint(*), cent(*)
      real(c_double), value :: conlen
      integer(c_int) :: id
    end function
    integer(c_int) function mcp_block_destroy(ctx, id) bind(C, name="mcp_block_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
    end function
    integer(c_int) function mcp_grid_create(ctx, mxlam, npts, r, ptab, dr, id) bind(C, name="mcp_grid_create")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mxlam, npts
      real(c_double) :: r(*), ptab(*)
      real(c_double), value :: dr
      integer(c_int) :: id
    end function
    integer(c_int) function mcp_grid_destroy(ctx, id) bind(C, name="mcp_grid_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
    end function
    integer(c_int) function mcp_grid_create_airy(ctx, mxlam, nstep, r, drnow, ptab, id) &
        bind(C, name="mcp_grid_create_airy")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mxlam, nstep
      real(c_double) :: r(*), drnow(*), ptab(*)
      integer(c_int) :: id
    end function
    integer(c_int) function mcp_hammat(ctx, block_id, grid_id, nr, idx, pconst, ered, w) bind(C, name="mcp_hammat")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, pconst
      integer(c_int), value :: block_id, grid_id, nr
      integer(c_int) :: idx(*)
      real(c_double), value :: ered
      real(c_double) :: w(*)
    end function
    integer(c_int) function mcp_yinit(ctx, block_id, grid_id, k0, outward, bcyo, ninst, ered, pconst, y) &
        bind(C, name="mcp_yinit")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, pconst
      integer(c_int), value :: block_id, grid_id, k0, outward, ninst
      real(c_double), value :: bcyo
      real(c_double) :: ered(*), y(*)
    end function
    integer(c_int) function mcp_prop_ld(ctx, block_id, grid_id, iprop, ninst, ered, pconst, y, nodes, status) &
        bind(C, name="mcp_prop_ld")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, pconst
      integer(c_int), value :: block_id, grid_id, iprop, ninst
      real(c_double) :: ered(*), y(*)
      integer(c_int) :: nodes(*), status(*)
    end function
    integer(c_int) function mcp_prop_ld_multi(ctx, nblk, block_ids, grid_id, ninst, ered, y, nodes, status) &
        bind(C, name="mcp_prop_ld_multi")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nblk, grid_id
      integer(c_int) :: block_ids(*), ninst(*), nodes(*), status(*)
      real(c_double) :: ered(*), y(*)
    end function
    integer(c_int) function mcp_set_graph(ctx, on) bind(C, name="mcp_set_graph")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function mcp_mg_scheme(imgsel, nsub, a, b) bind(C, name="mcp_mg_scheme")
      import :: c_int, c_double
      integer(c_int), value :: imgsel
      integer(c_int) :: nsub
      real(c_double) :: a(*), b(*)
    end function
    integer(c_int) function mcp_prop_mg(ctx, block_id, grid_id, imgsel, nstep, dr, ninst, ered, pconst, y, &
        nodes, status) bind(C, name="mcp_prop_mg")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, pconst
      integer(c_int), value :: block_id, grid_id, imgsel, nstep, ninst
      real(c_double), value :: dr
      real(c_double) :: ered(*), y(*)
      integer(c_int) :: nodes(*), status(*)
    end function
    integer(c_int) function mcp_prop_airy(ctx, block_id, grid_id, ninst, ered, y, nodes, status) &
        bind(C, name="mcp_prop_airy")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: block_id, grid_id, ninst
      real(c_double) :: ered(*), y(*)
      integer(c_int) :: nodes(*), status(*)
    end function
    integer(c_int) function mcp_block_set_potl(ctx, block_id, nterm, lam, npower, a, e, rscale, ep2ru) &
        bind(C, name="mcp_block_set_potl")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: block_id, nterm
      integer(c_int) :: lam(*), npower(*)
      real(c_double) :: a(*), e(*)
      real(c_double), value :: rscale, ep2ru
    end function
    integer(c_int) function mcp_prop_vivs(ctx, block_id, ninst, ered, rmat, rstart, rstop, drnow, drmax, &
        tldiag, toff, replay, max_steps, steps, nstep, status) bind(C, name="mcp_prop_vivs")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: block_id, ninst, replay, max_steps
      real(c_double) :: ered(*), rmat(*), steps(*)
      real(c_double), value :: rstart, rstop, drnow, drmax, tldiag, toff
      integer(c_int) :: nstep(*), status(*)
    end function
    integer(c_int) function mcp_prop_airy_tolhi(ctx, block_id, ninst, ered, z, rstart, rstop, dr0, powr, &
        tolhi, replay, max_steps, steps, nrec, nodes, nstep, status) bind(C, name="mcp_prop_airy_tolhi")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: block_id, ninst, replay, max_steps
      real(c_double) :: ered(*), z(*), steps(*)
      real(c_double), value :: rstart, rstop, dr0, powr, tolhi
      integer(c_int) :: nrec, nodes(*), nstep(*), status(*)
    end function
    integer(c_int) function mcp_ytrans(ctx, block_id, ninst, pconst, y, eint) bind(C, name="mcp_ytrans")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, pconst, y
      integer(c_int), value :: block_id, ninst
      real(c_double) :: eint(*)
    end function
    integer(c_int) function mcp_ytok_ktos(ctx, n, ninst, y, wvec, nopen, nb, cent, rval, kmat, sr, si, status) &
        bind(C, name="mcp_ytok_ktos")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, ninst
      real(c_double) :: y(*), wvec(*), cent(*), kmat(*), sr(*), si(*)
      integer(c_int) :: nopen(*), nb(*), status(*)
      real(c_double), value :: rval
    end function
    integer(c_int) function mcp_smatrix_ld(ctx, block_id, grid_id, iprop, ninst, ered, pconst, wvec, nopen, nb, &
        rval, kmat, sr, si, nodes, status) bind(C, name="mcp_smatrix_ld")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, pconst
      integer(c_int), value :: block_id, grid_id, iprop, ninst
      real(c_double) :: ered(*), wvec(*), kmat(*), sr(*), si(*)
      integer(c_int) :: nopen(*), nb(*), nodes(*), status(*)
      real(c_double), value :: rval
    end function
    integer(c_int) function mcp_smatrix_segments(ctx, block_id, nseg, grid_ids, iprops, ninst, ered, pconst, &
        wvec, nopen, nb, rval, kmat, sr, si, nodes, status) bind(C, name="mcp_smatrix_segments")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, pconst
      integer(c_int), value :: block_id, nseg, ninst
      integer(c_int) :: grid_ids(*), iprops(*), nopen(*), nb(*), nodes(*), status(*)
      real(c_double) :: ered(*), wvec(*), kmat(*), sr(*), si(*)
      real(c_double), value :: rval
    end function
    integer(c_int) function mcp_bound_match(ctx, block_id, grid_out, grid_in, iprop_out, iprop_in, ninst, ered, &
        pconst, nodec, eigmin, imin, psimch, evals, status) bind(C, name="mcp_bound_match")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, pconst, psimch, evals
      integer(c_int), value :: block_id, grid_out, grid_in, iprop_out, iprop_in, ninst
      real(c_double) :: ered(*), eigmin(*)
      integer(c_int) :: nodec(*), imin(*), status(*)
    end function
    integer(c_int) function mcp_set_team_size(ctx, g) bind(C, name="mcp_set_team_size")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: g
    end function
    integer(c_int) function mcp_set_team_launch(ctx, cooperative) bind(C, name="mcp_set_team_launch")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: cooperative
    end function
    integer(c_int) function mcp_set_large_mode(ctx, mode) bind(C, name="mcp_set_large_mode")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function
    integer(c_int) function mcp_set_inverse_kernel(ctx, mode) bind(C, name="mcp_set_inverse_kernel")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function
    integer(c_int) function mcp_bound_wave(ctx, block_id, gid_out, gid_in, ered, pconst, nodec, eigmin, psimch, &
        psi_out, psi_in, status) bind(C, name="mcp_bound_wave")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: block_id, gid_out, gid_in
      real(c_double), value :: ered
      real(c_double) :: pconst(*), eigmin, psimch(*), psi_out(*), psi_in(*)
      integer(c_int) :: nodec, status
    end function
    integer(c_int) function mcp_ytrans_cent(ctx, block_id, ninst, pconst, y, eint, cent) &
        bind(C, name="mcp_ytrans_cent")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: block_id, ninst
      real(c_double) :: pconst(*), y(*), eint(*), cent(*)
    end function
    integer(c_int) function mcp_set_degtol(ctx, degtol) bind(C, name="mcp_set_degtol")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: degtol
    end function
    integer(c_int) function mcp_set_vivs_options(ctx, iv, ivp, ivpp, ishift, idiag, ipert, isym, ialpha, ialfp, &
        alpha1, alpha2, xsqmax) bind(C, name="mcp_set_vivs_options")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: iv, ivp, ivpp, ishift, idiag, ipert, isym, ialpha, ialfp
      real(c_double), value :: alpha1, alpha2, xsqmax
    end function
    integer(c_int) function mcp_set_precision(ctx, newton_steps, r_fp64) bind(C, name="mcp_set_precision")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: newton_steps
      real(c_double), value :: r_fp64
    end function
    integer(c_int) function mcp_precision_stats(ctx, mixed_matrices, fp64_again) bind(C, name="mcp_precision_stats")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: mixed_matrices, fp64_again
    end function
    integer(c_int) function mcp_set_md_kernel(ctx, mode) bind(C, name="mcp_set_md_kernel")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function
    integer(c_int) function mcp_set_lds_threads(ctx, nt) bind(C, name="mcp_set_lds_threads")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt
    end function
    integer(c_int) function mcp_last_timing(ctx, ms_sector, ms_prop, ms_total) bind(C, name="mcp_last_timing")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: ms_sector, ms_prop, ms_total
    end function
    integer(c_int) function mcp_last_segment_timing(ctx, k, ms_sector, ms_prop) &
        bind(C, name="mcp_last_segment_timing")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
      real(c_double) :: ms_sector, ms_prop
    end function
  end interface

  ! ------------------------------------------------------------------
  ! state kept by the drop-in layer
  ! ------------------------------------------------------------------
  integer, parameter :: MCPF_NCACHE = 4

  type mcpf_block_entry
    integer(c_int) :: id = -1
    integer :: n = 0, mxlam = 0, nconst = 0, nrsq = 0
    real(c_double) :: conlen = 0.0d0
    real(c_double), allocatable :: vl(:), eint(:), cent(:)
    integer :: used = 0
  end type

  type mcpf_grid_entry
    integer(c_int) :: id = -1
    integer :: kind = 0            ! 1 = LD grid (mcp_grid_create), 2 = AIRY grid
    integer :: mxlam = 0, npts = 0
    real(c_double) :: dr = 0.0d0
    real(c_double), allocatable :: r(:), ptab(:), drnow(:)
    integer :: used = 0
  end type

  type(c_ptr), save :: mcpf_ctx = c_null_ptr
  type(mcpf_block_entry), save :: mcpf_blocks(MCPF_NCACHE)
  type(mcpf_grid_entry), save :: mcpf_grids(MCPF_NCACHE)
  integer, save :: mcpf_clock = 0
  integer, save :: mcpf_uploads_block = 0, mcpf_uploads_grid = 0

  ! &POTL terms for the propagators that choose their radii as they go
  ! (AIRY with TOLHI > 0, VIVS); registered by the host once it has read
  ! its potential (mcpf_set_potl).
  integer, save :: mcpf_potl_nterm = -1
  integer(c_int), allocatable, save :: mcpf_potl_lam(:), mcpf_potl_npower(:)
  real(c_double), allocatable, save :: mcpf_potl_a(:), mcpf_potl_e(:)
  ! SCAPOT's SCALAM (module efvs in the reference, source_code/scapot-default.f:
  ! 2-19): the factor on those radial terms, given with mcpf_set_scalam.
  real(c_double), save :: mcpf_scalam = 1.0d0

  ! VVPROP's steps as the energy that chose them recorded them (the part of
  ! the reference's scratch file ISCRU that the later energies read back):
  ! 4 values per step, mcpf_vivs_nrec steps.
  real(c_double), allocatable, save :: mcpf_vivs_rec(:)
  integer, save :: mcpf_vivs_nrec = 0

contains

  ! The library's message for the last failure on ctx, as a Fortran string.
  function mcpf_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    p = mcp_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function

  ! An entry point returned rc /= MCP_OK: print the routine, the code and the
  ! library's message on unit 6 and STOP, as the reference does on its own
  ! errors (e.g. source_code/mdprop.f:282-287).
  subroutine mcpf_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc == MCP_OK) return
    write (6, '(3A,I4,2A)') ' *** ERROR IN ', where, ': mcprop returned', rc, ': ', trim(mcpf_message(mcpf_ctx))
    stop
  end subroutine

  ! The context on the GPU (device 0, or the one MCPROP_DEVICE names),
  ! created on first use.
  function mcpf_context() result(ctx)
    type(c_ptr) :: ctx
    integer :: dev, ios, length
    character(len=16) :: val
    if (.not. c_associated(mcpf_ctx)) then
      dev = 0
      call get_environment_variable('MCPROP_DEVICE', val, length)
      if (length > 0) then
        read (val, *, iostat=ios) dev
        if (ios /= 0) dev = 0
      end if
      call mcpf_check(mcp_ctx_create(int(dev, c_int), mcpf_ctx), 'mcpf_context')
    end if
    ctx = mcpf_ctx
  end function

  integer function mcpf_lru_block() result(k)
    integer :: i
    k = 1
    do i = 1, MCPF_NCACHE
      if (mcpf_blocks(i)%id < 0) then
        k = i
        return
      end if
      if (mcpf_blocks(i)%used < mcpf_blocks(k)%used) k = i
    end do
  end function

  integer function mcpf_lru_grid() result(k)
    integer :: i
    k = 1
    do i = 1, MCPF_NCACHE
      if (mcpf_grids(i)%id < 0) then
        k = i
        return
      end if
      if (mcpf_grids(i)%used < mcpf_grids(k)%used) k = i
    end do
  end function

  ! Device id of the block (VL(NHAM, N(N+1)/2) with NHAM = mxlam + nconst +
  ! nrsq, EINT, CENT, CONLEN): an entry with the same data is reused, else
  ! the least recently used entry is replaced.
  function mcpf_block(n, mxlam, nconst, nrsq, vl, eint, cent, conlen) result(id)
    integer, intent(in) :: n, mxlam, nconst, nrsq
    real(c_double), intent(in) :: vl(:), eint(:), cent(:)
    real(c_double), intent(in) :: conlen
    integer(c_int) :: id
    type(c_ptr) :: ctx
    integer :: i, k
    ctx = mcpf_context()
    mcpf_clock = mcpf_clock + 1
    do i = 1, MCPF_NCACHE
      associate (b => mcpf_blocks(i))
        if (b%id < 0) cycle
        if (b%n /= n .or. b%mxlam /= mxlam .or. b%nconst /= nconst .or. b%nrsq /= nrsq) cycle
        if (b%conlen /= conlen) cycle
        if (size(b%vl) /= size(vl)) cycle
        if (any(b%vl /= vl) .or. any(b%eint /= eint(1:n)) .or. any(b%cent /= cent(1:n))) cycle
        b%used = mcpf_clock
        id = b%id
        return
      end associate
    end do
    k = mcpf_lru_block()
    associate (b => mcpf_blocks(k))
      if (b%id >= 0) call mcpf_check(mcp_block_destroy(ctx, b%id), 'mcpf_block')
      b%id = -1
      call mcpf_check(mcp_block_create(ctx, int(n, c_int), int(mxlam, c_int), int(nconst, c_int), &
                                       int(nrsq, c_int), vl, eint, cent, conlen, id), 'mcpf_block')
      b%id = id
      b%n = n
      b%mxlam = mxlam
      b%nconst = nconst
      b%nrsq = nrsq
      b%conlen = conlen
      b%vl = vl
      b%eint = eint(1:n)
      b%cent = cent(1:n)
      b%used = mcpf_clock
    end associate
    mcpf_uploads_block = mcpf_uploads_block + 1
  end function

  ! Device id of a radial grid: kind 1 = LD grid (npts radii, P table
  ! ptab(mxlam, npts), step dr), kind 2 = AIRY grid (npts = 2 nstep + 1 radii,
  ! drnow(nstep)).  Reused when the same grid and table were uploaded before.
  function mcpf_grid(kind, mxlam, npts, r, ptab, dr, drnow) result(id)
    integer, intent(in) :: kind, mxlam, npts
    real(c_double), intent(in) :: r(:), ptab(:), dr, drnow(:)
    integer(c_int) :: id
    type(c_ptr) :: ctx
    integer :: i, k
    ctx = mcpf_context()
    mcpf_clock = mcpf_clock + 1
    do i = 1, MCPF_NCACHE
      associate (g => mcpf_grids(i))
        if (g%id < 0) cycle
        if (g%kind /= kind .or. g%mxlam /= mxlam .or. g%npts /= npts .or. g%dr /= dr) cycle
        if (any(g%r /= r(1:npts)) .or. any(g%ptab /= ptab(1:mxlam*npts))) cycle
        if (kind == 2) then
          if (any(g%drnow /= drnow(1:(npts - 1)/2))) cycle
        end if
        g%used = mcpf_clock
        id = g%id
        return
      end associate
    end do
    k = mcpf_lru_grid()
    associate (g => mcpf_grids(k))
      if (g%id >= 0) call mcpf_check(mcp_grid_destroy(ctx, g%id), 'mcpf_grid')
      g%id = -1
      if (kind == 1) then
        call mcpf_check(mcp_grid_create(ctx, int(mxlam, c_int), int(npts, c_int), r, ptab, dr, id), 'mcpf_grid')
        g%drnow = [real(c_double) ::]
      else
        call mcpf_check(mcp_grid_create_airy(ctx, int(mxlam, c_int), int((npts - 1)/2, c_int), r, drnow, ptab, &
                                             id), 'mcpf_grid')
        g%drnow = drnow(1:(npts - 1)/2)
      end if
      g%id = id
      g%kind = kind
      g%mxlam = mxlam
      g%npts = npts
      g%dr = dr
      g%r = r(1:npts)
      g%ptab = ptab(1:mxlam*npts)
      g%used = mcpf_clock
    end associate
    mcpf_uploads_grid = mcpf_uploads_grid + 1
  end function

  ! Device objects held by the cache and uploads so far (a host program or a
  ! test can check that repeated calls do not grow the device state).
  subroutine mcpf_cache_stats(nblocks, ngrids, nup_block, nup_grid)
    integer, intent(out) :: nblocks, ngrids, nup_block, nup_grid
    nblocks = count(mcpf_blocks(:)%id >= 0)
    ngrids = count(mcpf_grids(:)%id >= 0)
    nup_block = mcpf_uploads_block
    nup_grid = mcpf_uploads_grid
  end subroutine

  ! Register the &POTL radial terms (POTENL case 1, source_code/potenl.f:
  ! 160-200): term t adds a(t)*R**npower(t) (npower /= 0) or a(t)*exp(e(t)*R)
  ! to radial coefficient lam(t) (1-based, <= MXLAM).  Needed only by the
  ! propagators that evaluate the potential where their step control lands.
  subroutine mcpf_set_potl(nterm, lam, npower, a, e)
    integer, intent(in) :: nterm
    integer, intent(in) :: lam(nterm), npower(nterm)
    real(c_double), intent(in) :: a(nterm), e(nterm)
    mcpf_potl_nterm = nterm
    mcpf_potl_lam = int(lam - 1, c_int)
    mcpf_potl_npower = int(npower, c_int)
    mcpf_potl_a = a
    mcpf_potl_e = e
  end subroutine

  ! SCALAM for the &POTL terms above (the caller's efvs SCALAM).
  subroutine mcpf_set_scalam(scalam)
    real(c_double), intent(in) :: scalam
    mcpf_scalam = scalam
  end subroutine

  ! Destroy every cached device object and the context.
  subroutine mcpf_release()
    integer :: i
    integer(c_int) :: rc
    if (.not. c_associated(mcpf_ctx)) return
    do i = 1, MCPF_NCACHE
      if (mcpf_blocks(i)%id >= 0) rc = mcp_block_destroy(mcpf_ctx, mcpf_blocks(i)%id)
      mcpf_blocks(i)%id = -1
      if (mcpf_grids(i)%id >= 0) rc = mcp_grid_destroy(mcpf_ctx, mcpf_grids(i)%id)
      mcpf_grids(i)%id = -1
    end do
    call mcp_ctx_destroy(mcpf_ctx)
    mcpf_ctx = c_null_ptr
  end subroutine

end module mcprop
