! Fortran side of the on-disk format checks.  The WRITE and READ statements
! are those of the reference's writers and readers (writewf and
! write_wf_for_sgw_debug, src/write_wf.f90; read_wf_tddft,
! src/read_wf.f90:130-228; eig_bar, src/diag.f90:109-115; plot_dip,
! src/tddft_plot.f90:100-105; write_to_restart / read_phi_bar,
! src/tddft_write_to_restart.f90, src/read_phi_bar.f90), so the files this
! program writes are the reference's files for the same data.
!
!   formats_check write      in.raw -> wf_long.bin wf_short.bin wf_long.txt
!                                      wf_short.txt eig_f d_dip_f.dat
!                                      phi_f.bin info_f
!   formats_check read <wf file> <bin|txt> <N> <nstates-max>
!                                   -> read.raw (what read_wf_tddft read)
!   formats_check eig <file> <n>    -> eig.raw
!   formats_check dip <file> <nrow> <ncol> -> dip.raw
!   formats_check info <file>       -> info.raw
!   formats_check rec <file> <N> <nrec> -> rec.raw (complex*16 records)
program formats_check
  implicit none
  character(len=256) :: cmd, a1, a2, a3, a4
  call get_command_argument(1, cmd)
  select case (trim(cmd))
  case ('write')
    call do_write()
  case ('read')
    call get_command_argument(2, a1); call get_command_argument(3, a2)
    call get_command_argument(4, a3); call get_command_argument(5, a4)
    call do_read(trim(a1), trim(a2) == 'bin', iarg(a3), iarg(a4))
  case ('eig')
    call get_command_argument(2, a1); call get_command_argument(3, a2)
    call do_eig(trim(a1), iarg(a2))
  case ('dip')
    call get_command_argument(2, a1); call get_command_argument(3, a2); call get_command_argument(4, a3)
    call do_dip(trim(a1), iarg(a2), iarg(a3))
  case ('info')
    call get_command_argument(2, a1)
    call do_info(trim(a1))
  case ('rec')
    call get_command_argument(2, a1); call get_command_argument(3, a2); call get_command_argument(4, a3)
    call do_rec(trim(a1), iarg(a2), iarg(a3))
  case default
    stop 'unknown command'
  end select

contains

  integer function iarg(s)
    character(len=*), intent(in) :: s
    read(s, *) iarg
  end function

  subroutine do_write()
    integer :: nx, ny, nz, ns, nrow, nb, nds, it, i, k
    real*8 :: dx, dy, dz
    real*8, allocatable :: eig(:), tmp(:, :), dd(:, :)
    complex*16, allocatable :: z(:)
    integer :: form
    open(10, file='in.raw', access='stream', form='unformatted', status='old')
    read(10) nx, ny, nz, ns, nrow, nb, nds, it
    read(10) dx, dy, dz
    allocate(eig(ns), tmp(nx*ny*nz, ns), dd(4, nrow), z(nx*ny*nz))
    read(10) eig, tmp, dd
    close(10)
    ! the stored orbital of the reference's writers: sign(Re phi)|phi| of the
    ! complex orbital phi, here phi(k) = tmp(k) + i tmp(k+1)/2 (cyclic)
    do i = 1, ns
      do k = 1, nx*ny*nz
        z(k) = cmplx(tmp(k, i), 0.5d0*tmp(mod(k, nx*ny*nz) + 1, i), kind=8)
      end do
      do k = 1, nx*ny*nz
        if (dble(z(k)) > 0d0) then
          tmp(k, i) = abs(z(k))
        else
          tmp(k, i) = -1d0*abs(z(k))
        end if
      end do
    end do
    do form = 0, 1
      if (form == 0) then
        open(unit=1, file='wf_long.bin', form='unformatted')
        write(1) 'nx      ', nx
        write(1) 'ny      ', ny
        write(1) 'nz      ', nz
        write(1) 'dx      ', dx
        write(1) 'dy      ', dy
        write(1) 'dz      ', dz
        write(1) 'nsp     ', 1
        write(1) 'nstates ', ns
        write(1) 'orbmn_wf', 1
        write(1) 'ihomo   ', 2
        write(1) 'evls    '
        write(1) eig
        write(1) 'orbitals'
        do i = 1, ns
          write(1) i, 1
          write(1) tmp(:, i)
        end do
        close(1)
        open(unit=1, file='wf_short.bin', form='unformatted')
        write(1) 'nx      ', nx
        write(1) 'ny      ', ny
        write(1) 'nz      ', nz
        write(1) 'dx      ', dx
        write(1) 'dy      ', dy
        write(1) 'dz      ', dz
        write(1) 'nsp     ', 1
        write(1) 'nstates ', ns
        write(1) 'evls    '
        write(1) eig
        write(1) 'orbitals'
        do i = 1, ns
          write(1) i, 1
          write(1) tmp(:, i)
        end do
        close(1)
      else
        open(unit=1, file='wf_long.txt')
        write(1, *) 'nx      ', nx
        write(1, *) 'ny      ', ny
        write(1, *) 'nz      ', nz
        write(1, *) 'dx      ', dx
        write(1, *) 'dy      ', dy
        write(1, *) 'dz      ', dz
        write(1, *) 'nsp     ', 1
        write(1, *) 'nstates ', ns
        write(1, *) 'orbmn_wf', 1
        write(1, *) 'ihomo   ', 2
        write(1, *) 'evls    '
        write(1, *) eig
        write(1, *) 'orbitals'
        do i = 1, ns
          write(1, *) i, 1
          write(1, *) tmp(:, i)
        end do
        close(1)
        open(unit=1, file='wf_short.txt')
        write(1, *) 'nx      ', nx
        write(1, *) 'ny      ', ny
        write(1, *) 'nz      ', nz
        write(1, *) 'dx      ', dx
        write(1, *) 'dy      ', dy
        write(1, *) 'dz      ', dz
        write(1, *) 'nsp     ', 1
        write(1, *) 'nstates ', ns
        write(1, *) 'evls    '
        write(1, *) eig
        write(1, *) 'orbitals'
        do i = 1, ns
          write(1, *) i, 1
          write(1, *) tmp(:, i)
        end do
        close(1)
      end if
    end do
    open(unit=1, file='eig_f')
    write(1, *) eig
    close(1)
    open(unit=15, file='d_dip_f.dat')
    do k = 1, nrow
      write(15, 888) dd(1, k), dd(2:4, k)
    end do
    close(15)
888 format (' ',4e18.8)
    open(unit=305, file='info_f')
    write(305, *) nb, nds, it
    close(305)
    open(3000, file='phi_f.bin', form='unformatted')
    do i = 1, ns
      z = cmplx(tmp(:, i), -0.5d0 * tmp(:, i) + dble(i), kind=8)
      write(3000) z
    end do
    close(3000)
  end subroutine

  subroutine do_read(fn, bin, n, nmax)
    character(len=*), intent(in) :: fn
    logical, intent(in) :: bin
    integer, intent(in) :: n, nmax
    character*8 ch
    integer :: nx_r, ny_r, nz_r, nsp_r, nstates_r, i_orb, i
    real*8 :: dx_r, dy_r, dz_r
    real*8 :: tmp(n)
    if (bin) then
      open(441, file=fn, status='old', form='unformatted')
      read(441) ch, nx_r
      read(441) ch, ny_r
      read(441) ch, nz_r
      read(441) ch, dx_r
      read(441) ch, dy_r
      read(441) ch, dz_r
      read(441) ch, nsp_r
      read(441) ch, nstates_r
      read(441) ch
      read(441)
      read(441) ch
    else
      open(441, file=fn, status='old')
      read(441, *) ch, nx_r
      read(441, *) ch, ny_r
      read(441, *) ch, nz_r
      read(441, *) ch, dx_r
      read(441, *) ch, dy_r
      read(441, *) ch, dz_r
      read(441, *) ch, nsp_r
      read(441, *) ch, nstates_r
      read(441, *) ch
      read(441, *)
      read(441, *) ch
    end if
    if (nstates_r > nmax) stop 'too many states'
    open(20, file='read.raw', access='stream', form='unformatted', status='replace')
    write(20) nx_r, ny_r, nz_r, nsp_r, nstates_r, dx_r, dy_r, dz_r
    do i = 1, nstates_r
      if (bin) then
        read(441) i_orb, nsp_r
        read(441) tmp
      else
        read(441, *) i_orb, nsp_r
        read(441, *) tmp
      end if
      write(20) i_orb, nsp_r, tmp
    end do
    close(20)
    close(441)
  end subroutine

  subroutine do_eig(fn, n)
    character(len=*), intent(in) :: fn
    integer, intent(in) :: n
    real*8 :: eig(n)
    open(unit=1, file=fn, status='old')
    read(1, *) eig
    close(1)
    open(20, file='eig.raw', access='stream', form='unformatted', status='replace')
    write(20) eig
    close(20)
  end subroutine

  subroutine do_dip(fn, nrow, ncol)
    character(len=*), intent(in) :: fn
    integer, intent(in) :: nrow, ncol
    real*8 :: row(ncol)
    integer :: k
    open(unit=16, file=fn, status='old')
    open(20, file='dip.raw', access='stream', form='unformatted', status='replace')
    do k = 1, nrow
      read(16, *) row
      write(20) row
    end do
    close(20)
    close(16)
  end subroutine

  subroutine do_info(fn)
    character(len=*), intent(in) :: fn
    integer :: nb, nds, it
    open(unit=34, file=fn, status='old')
    read(34, *) nb, nds, it
    close(34)
    open(20, file='info.raw', access='stream', form='unformatted', status='replace')
    write(20) nb, nds, it
    close(20)
  end subroutine

  subroutine do_rec(fn, n, nrec)
    character(len=*), intent(in) :: fn
    integer, intent(in) :: n, nrec
    complex*16 :: z(n)
    integer :: i
    open(234, file=fn, status='old', form='unformatted')
    open(20, file='rec.raw', access='stream', form='unformatted', status='replace')
    do i = 1, nrec
      read(234) z
      write(20) z
    end do
    close(20)
    close(234)
  end subroutine

end program
