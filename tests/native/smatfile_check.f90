! Fortran side of the S-matrix file tests: unformatted sequential I/O of
! the record layout that SPROC writes (source_code/sproc.f:254-281,
! 797-807) and HDREAD/SLPRD read (source_code/sread.f), done by the
! Fortran compiler's own runtime.
!   smatfile_check read  FILE   prints the header and every entry as text
!   smatfile_check write FILE   writes a fixed file of two entries with
!                               two external field variables
program smatfile_check
  implicit none
  character(len=512) :: mode, path
  call get_command_argument(1, mode)
  call get_command_argument(2, path)
  if (trim(mode) == 'read') then
    call read_file(trim(path))
  else if (trim(mode) == 'write') then
    call write_file(trim(path))
  else
    write(*, '(A)') 'usage: smatfile_check read|write FILE'
    stop 2
  end if

contains

  subroutine read_file(path)
    character(*), intent(in) :: path
    character(80) :: label
    character(20) :: efvnam(10)
    character(6) :: efvunt(10)
    integer :: itype, nstate, nqn, iprogm, nlvl, ndgvl, nconst, nrsq, ibound, itpsub
    integer :: nefv, isvefv, nefvp, nfield, nnrg, i, j, ios
    integer :: jtot, inrg, m, ifield, iexch, nopen
    double precision :: ured, energn, eref, wt
    integer, allocatable :: jstate(:), lev(:), lv(:)
    double precision, allocatable :: elevel(:), energy(:), efv(:), wvec(:), sr(:, :), si(:, :)

    open(10, file=path, form='unformatted', access='sequential', status='old')
    read(10) label, itype, nstate, nqn, ured, iprogm
    allocate(jstate(nstate * nqn))
    read(10) (jstate(i), i = 1, nstate * nqn)
    read(10) nlvl
    backspace(10)
    allocate(elevel(nlvl))
    read(10) nlvl, (elevel(i), i = 1, nlvl)
    read(10) ndgvl, nconst, nrsq, ibound, itpsub
    read(10) nefv, isvefv
    backspace(10)
    read(10) nefv, isvefv, (efvnam(i), efvunt(i), i = 1, nefv)
    nefvp = 0
    if (nefv > 0) nefvp = max(nefv, isvefv)
    read(10) nfield, nnrg
    backspace(10)
    allocate(energy(nnrg))
    read(10) nfield, nnrg, (energy(i), i = 1, nnrg)
    write(*, '(A,A)') 'LABEL ', trim(label)
    write(*, '(A,4I8)') 'HEAD ', itype, nstate, nqn, iprogm
    write(*, '(A,ES25.16E3)') 'URED ', ured
    write(*, '(A,100I6)') 'JSTATE ', jstate
    write(*, '(A,I6)') 'NLEVEL ', nlvl
    do i = 1, nlvl
      write(*, '(A,ES25.16E3)') 'ELEVEL ', elevel(i)
    end do
    write(*, '(A,5I6)') 'FLAGS ', ndgvl, nconst, nrsq, ibound, itpsub
    write(*, '(A,2I6)') 'EFV ', nefv, isvefv
    do i = 1, nefv
      write(*, '(A,A,A,A)') 'EFVNAME ', efvnam(i), '|', efvunt(i)
    end do
    write(*, '(A,2I6)') 'NRG ', nfield, nnrg
    do i = 1, nnrg
      write(*, '(A,ES25.16E3)') 'ENERGY ', energy(i)
    end do
    allocate(efv(0:nefvp))
    do
      read(10, iostat=ios) jtot, inrg, m, ifield, energn, (efv(i), i = 0, nefvp), eref, iexch, wt, nopen
      if (ios /= 0) exit
      allocate(lev(nopen), lv(nopen), wvec(nopen), sr(nopen, nopen), si(nopen, nopen))
      read(10) (lev(i), lv(i), wvec(i), i = 1, nopen)
      read(10) ((sr(i, j), j = 1, i), i = 1, nopen)
      read(10) ((si(i, j), j = 1, i), i = 1, nopen)
      write(*, '(A,6I6)') 'ENTRY ', jtot, inrg, m, ifield, iexch, nopen
      write(*, '(A,3ES25.16E3)') 'REAL ', energn, eref, wt
      do i = 0, nefvp
        write(*, '(A,ES25.16E3)') 'EFVVAL ', efv(i)
      end do
      do i = 1, nopen
        write(*, '(A,2I6,ES25.16E3)') 'CHAN ', lev(i), lv(i), wvec(i)
      end do
      do i = 1, nopen
        do j = 1, i
          write(*, '(A,2I6,2ES25.16E3)') 'S ', i, j, sr(i, j), si(i, j)
        end do
      end do
      deallocate(lev, lv, wvec, sr, si)
    end do
    close(10)
  end subroutine read_file

  subroutine write_file(path)
    character(*), intent(in) :: path
    character(80) :: label
    character(20) :: efvnam(2)
    character(6) :: efvunt(2)
    integer :: jstate(6), i, j, k, nopen
    double precision :: elevel(3), energy(2), s(3, 3)

    label = 'written by the Fortran runtime'
    jstate = (/ 0, 2, 4, 1, 2, 3 /)
    elevel = (/ 0.d0, 180.d0, 600.d0 /)
    energy = (/ 1000.d0, 900.d0 /)
    efvnam(1) = 'MAGNETIC Z FIELD'
    efvunt(1) = 'GAUSS'
    efvnam(2) = 'ELECTRIC FIELD'
    efvunt(2) = 'KV/CM'
    open(11, file=path, form='unformatted', access='sequential', status='replace')
    write(11) label, 1, 3, 2, 20.d0, 19
    write(11) (jstate(i), i = 1, 6)
    write(11) 3, (elevel(i), i = 1, 3)
    write(11) 0, 0, 0, 0, 0
    write(11) 2, 1, (efvnam(i), efvunt(i), i = 1, 2)
    write(11) 1, 2, (energy(i), i = 1, 2)
    do k = 1, 2
      nopen = k + 1
      do i = 1, nopen
        do j = 1, nopen
          s(i, j) = 0.125d0 * dble(k) + 0.01d0 * dble(i * j) + 1.d-3 * dble(i + j)
        end do
      end do
      write(11) 10, k, 1, 1, energy(k), 0.d0, 250.d0 * dble(k), 1.5d0, -3.25d0, 0, 1.d0, nopen
      write(11) (i, 2 * i - 1, 0.5d0 * dble(i) + 0.25d0 * dble(k), i = 1, nopen)
      write(11) ((s(i, j), j = 1, i), i = 1, nopen)
      write(11) ((-s(i, j), j = 1, i), i = 1, nopen)
    end do
    close(11)
  end subroutine write_file

end program smatfile_check
