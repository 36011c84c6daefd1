! iso_c_binding interfaces to libopaw_hip.so (include/opaw_hip.h) and the
! wrappers a Fortran host calls in place of the reference's module-state
! routines (the operators and rk4_prop_shs under their own names are in
! opaw_dropin.f90).  The module keeps one context (the device copy of
! main_mod / atom_mod / tddft_mod state); arrays cross the boundary in the
! reference layouts (column-major, complex*16 interleaved).
module opaw_hip_mod
  use iso_c_binding
  implicit none
  private

  type(c_ptr), save, public :: opaw_ctx = c_null_ptr
  real(c_double), save :: opaw_dt = 0d0     ! tddft_mod dt, as handed over by gpu_set_dt

  interface
    integer(c_int) function opaw_create(ctx, nx, ny, nz, dx, dy, dz, periodic, precision, nsys, nb) &
        bind(C, name="opaw_create")
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: ctx
      integer(c_int), value :: nx, ny, nz, periodic, precision, nsys, nb
      real(c_double), value :: dx, dy, dz
    end function
    integer(c_int) function opaw_destroy(ctx) bind(C, name="opaw_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function opaw_ek_prep(ctx, ekcut) bind(C, name="opaw_ek_prep")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: ekcut
    end function
    integer(c_int) function opaw_set_atoms(ctx, natom, atom_map, ir_start, ntype, nrough, mstates, nl2) &
        bind(C, name="opaw_set_atoms")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: natom, ntype, nl2
      integer(c_int) :: atom_map(*), ir_start(*), nrough(*), mstates(*)
    end function
    integer(c_int) function opaw_set_projectors(ctx, ia, p3d, p3d1, s, sinv) bind(C, name="opaw_set_projectors")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ia
      real(c_double) :: p3d(*), p3d1(*), s(*), sinv(*)
    end function
    integer(c_int) function opaw_set_sphere(ctx, ia, ngrid, local_grid, local_g3d) bind(C, name="opaw_set_sphere")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ia, ngrid
      integer(c_int) :: local_grid(*)
      real(c_double) :: local_g3d(*)
    end function
    integer(c_int) function opaw_set_type(ctx, it, qijlm, dij0, eijkl) bind(C, name="opaw_set_type")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: it
      real(c_double) :: qijlm(*), dij0(*), eijkl(*)
    end function
    integer(c_int) function opaw_finalize_tables(ctx) bind(C, name="opaw_finalize_tables")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function opaw_set_potentials(ctx, vloc_tot, ncoret, rnel, usexcnhat) &
        bind(C, name="opaw_set_potentials")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: vloc_tot(*), ncoret(*)
      real(c_double), value :: rnel
      integer(c_int), value :: usexcnhat
    end function
    integer(c_int) function opaw_vk_prep(ctx) bind(C, name="opaw_vk_prep")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function opaw_write_orbitals(ctx, sys, host) bind(C, name="opaw_write_orbitals")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: sys
    end function
    integer(c_int) function opaw_read_orbitals(ctx, sys, host) bind(C, name="opaw_read_orbitals")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: sys
    end function
    integer(c_int) function opaw_make_hamiltonian_tddft(ctx, sys) bind(C, name="opaw_make_hamiltonian_tddft")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: sys
    end function
    integer(c_int) function opaw_rk4_prop_shs(ctx, dt) bind(C, name="opaw_rk4_prop_shs")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function opaw_plot_dip(ctx, sm, d_dip, dip, dip_pert) bind(C, name="opaw_plot_dip")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: sm
      real(c_double) :: d_dip(3), dip(3), dip_pert(3)
    end function
    integer(c_int) function opaw_excite_psi_bar_pert(ctx, from_sys, to_sys, sm, ipol) &
        bind(C, name="opaw_excite_psi_bar_pert")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: from_sys, to_sys, ipol
      real(c_double), value :: sm
    end function
    integer(c_int) function opaw_write_restart(ctx, dir, it) bind(C, name="opaw_write_restart")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: dir(*)
      integer(c_int), value :: it
    end function
    integer(c_int) function opaw_read_restart(ctx, dir, it) bind(C, name="opaw_read_restart")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: dir(*)
      integer(c_int) :: it
    end function
    integer(c_int) function opaw_set_nocc(ctx, nocc) bind(C, name="opaw_set_nocc")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nocc
    end function
    integer(c_int) function opaw_set_diis(ctx, ndi, mix) bind(C, name="opaw_set_diis")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ndi
      real(c_double), value :: mix
    end function
    integer(c_int) function opaw_ncpaw_make_hamiltonian(ctx, sys, use_diis) &
        bind(C, name="opaw_ncpaw_make_hamiltonian")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: sys, use_diis
    end function
    integer(c_int) function opaw_hminmax(ctx, sys, h_type, niter, fctr, seed, hmin, hmax) &
        bind(C, name="opaw_hminmax")
      import :: c_ptr, c_int, c_double, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: sys, h_type, niter
      real(c_double), value :: fctr
      integer(c_long_long), value :: seed
      real(c_double) :: hmin, hmax
    end function
    integer(c_int) function opaw_chebyshev_filter(ctx, sys, h_type, degree, hbot, hmax) &
        bind(C, name="opaw_chebyshev_filter")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: sys, h_type, degree
      real(c_double), value :: hbot, hmax
    end function
    integer(c_int) function opaw_orthog_phi(ctx, sys, h_type, eig) bind(C, name="opaw_orthog_phi")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: sys, h_type
      real(c_double) :: eig(*)
    end function
    integer(c_int) function opaw_diagh(ctx, sys, h_type, eig) bind(C, name="opaw_diagh")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: sys, h_type
      real(c_double) :: eig(*)
    end function
    integer(c_int) function opaw_h_phi_host(ctx, sys, pin, hp, nbatch) bind(C, name="opaw_h_phi_host")
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: sys, nbatch
      complex(c_double_complex) :: pin(*), hp(*)
    end function
    integer(c_int) function opaw_shs_host(ctx, sys, pin, pout, nbatch) bind(C, name="opaw_shs_host")
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: sys, nbatch
      complex(c_double_complex) :: pin(*), pout(*)
    end function
    integer(c_int) function opaw_sn_phi_host(ctx, pin, sp, nbatch, n) bind(C, name="opaw_sn_phi_host")
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbatch
      real(c_double), value :: n
      complex(c_double_complex) :: pin(*), sp(*)
    end function
    integer(c_int) function opaw_proj_host(ctx, ia, pin, which_p1, ca, ms) bind(C, name="opaw_proj_host")
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: ia, which_p1, ms
      complex(c_double_complex) :: pin(*), ca(*)
    end function
    integer(c_int) function opaw_write_orbitals_c16(ctx, sys, host) bind(C, name="opaw_write_orbitals_c16")
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: sys
      complex(c_double_complex) :: host(*)
    end function
    integer(c_int) function opaw_read_orbitals_c16(ctx, sys, host) bind(C, name="opaw_read_orbitals_c16")
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: sys
      complex(c_double_complex) :: host(*)
    end function
    integer(c_int) function opaw_write_dipoles(d_dip_path, dip_path, append, it0, nt, dt, d_dip, dip, nsys) &
        bind(C, name="opaw_write_dipoles")
      import :: c_int, c_double, c_char
      character(kind=c_char) :: d_dip_path(*), dip_path(*)
      integer(c_int), value :: append, it0, nt, nsys
      real(c_double), value :: dt
      real(c_double) :: d_dip(*), dip(*)
    end function
    integer(c_int) function opaw_set_dt(ctx, dt) bind(C, name="opaw_set_dt")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function opaw_rk4_prop_shs_host(ctx, p, pp, n, nb) bind(C, name="opaw_rk4_prop_shs_host")
      import :: c_ptr, c_int, c_long_long, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: p(*), pp(*)
      integer(c_long_long), value :: n
      integer(c_int), value :: nb
    end function
    function opaw_last_error() bind(C, name="opaw_last_error")
      import :: c_ptr
      type(c_ptr) :: opaw_last_error
    end function
  end interface

  public :: opaw_create, opaw_destroy, opaw_ek_prep, opaw_set_atoms, opaw_set_projectors, opaw_set_sphere, &
            opaw_set_type, opaw_finalize_tables, opaw_set_potentials, opaw_vk_prep, opaw_write_orbitals, &
            opaw_read_orbitals, opaw_make_hamiltonian_tddft, opaw_rk4_prop_shs, opaw_plot_dip, &
            opaw_excite_psi_bar_pert, opaw_write_restart, opaw_read_restart, opaw_set_nocc, opaw_set_diis, &
            opaw_ncpaw_make_hamiltonian, opaw_hminmax, opaw_chebyshev_filter, opaw_orthog_phi, opaw_diagh, &
            opaw_h_phi_host, opaw_shs_host, opaw_sn_phi_host, opaw_proj_host, opaw_write_orbitals_c16, &
            opaw_read_orbitals_c16, opaw_set_dt, opaw_rk4_prop_shs_host, opaw_write_dipoles, opaw_check, &
            opaw_require_ik
  public :: gpu_set_dt, gpu_scatter_phi_bar, gpu_gather_phi_bar, gpu_ncpaw_make_hamiltonian_tddft, gpu_plot_dip, &
            gpu_write_to_restart, gpu_read_phi_bar, gpu_ncpaw_make_hamiltonian, gpu_get_hminmax, gpu_chebyr_m, &
            gpu_orthog_phi, gpu_diagh

contains

  ! reference convention: any failure stops the rank (src/lib/library_lib2.f:338)
  subroutine opaw_check(stat, what)
    integer(c_int), intent(in) :: stat
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (stat /= 0) then
      call c_f_pointer(opaw_last_error(), msg, [512])
      n = 0
      do while (n < 512)
        if (msg(n + 1) == c_null_char) exit
        n = n + 1
      end do
      write(6, *) 'opaw_hip: ', what, ' failed: ', msg(1:n)
      error stop 'opaw_hip failure'
    end if
  end subroutine

  ! the context holds one kinetic table: the reference's k index ik is 1
  subroutine opaw_require_ik(ik, what)
    integer, intent(in) :: ik
    character(len=*), intent(in) :: what
    if (ik /= 1) then
      write(6, *) 'opaw_hip: ', what, ': only ik = 1 (one kinetic table), got ', ik
      error stop 'opaw_hip failure'
    end if
  end subroutine

  ! tddft_mod's dt, read by the rk4_prop_shs drop-in
  subroutine gpu_set_dt(dt)
    real(c_double), intent(in) :: dt
    call opaw_check(opaw_set_dt(opaw_ctx, dt), 'set_dt')
    opaw_dt = dt
  end subroutine

  ! scatter_phi_bar (src/tddft_scatter.f90): phi_bar and phi_bar_pert
  ! (n points x nb orbitals, complex*16) become the resident orbitals of
  ! systems 0 and 1
  subroutine gpu_scatter_phi_bar(p, pp, n, nb)
    integer, intent(in) :: n, nb
    complex(c_double_complex) :: p(n, nb), pp(n, nb)
    call opaw_check(opaw_write_orbitals_c16(opaw_ctx, 0_c_int, p), 'scatter phi_bar')
    call opaw_check(opaw_write_orbitals_c16(opaw_ctx, 1_c_int, pp), 'scatter phi_bar_pert')
  end subroutine

  ! gather_phi_bar: the resident orbitals back into phi_bar / phi_bar_pert
  subroutine gpu_gather_phi_bar(p, pp, n, nb)
    integer, intent(in) :: n, nb
    complex(c_double_complex) :: p(n, nb), pp(n, nb)
    call opaw_check(opaw_read_orbitals_c16(opaw_ctx, 0_c_int, p), 'gather phi_bar')
    call opaw_check(opaw_read_orbitals_c16(opaw_ctx, 1_c_int, pp), 'gather phi_bar_pert')
  end subroutine

  ! drop-in for ncpaw_make_hamiltonian_tddft (src/ncpaw_make_hamiltonian.f90:118)
  subroutine gpu_ncpaw_make_hamiltonian_tddft()
    call opaw_check(opaw_make_hamiltonian_tddft(opaw_ctx, -1_c_int), 'make_hamiltonian_tddft')
  end subroutine

  ! drop-in for plot_dip (src/tddft_plot.f90:1-105): the dipoles of both
  ! systems, and with the time step it given, the lines plot_dip writes at
  ! (it-1)*dt: d_dip.dat (' ',4e18.8) and dip.dat (list-directed t, dip,
  ! dip_pert); the files start afresh at it = 1, as the reference rewinds
  ! them there, and are appended to after (dt from gpu_set_dt)
  subroutine gpu_plot_dip(sm, d_dip, dip, dip_pert, it)
    real(c_double), intent(in) :: sm
    real(c_double) :: d_dip(3), dip(3), dip_pert(3)
    integer, intent(in), optional :: it
    real(c_double) :: both(6)
    integer(c_int) :: app
    call opaw_check(opaw_plot_dip(opaw_ctx, sm, d_dip, dip, dip_pert), 'plot_dip')
    if (present(it)) then
      both(1:3) = dip
      both(4:6) = dip_pert
      app = 1
      if (it == 1) app = 0
      call opaw_check(opaw_write_dipoles('d_dip.dat'//c_null_char, 'dip.dat'//c_null_char, app, &
                                         int(it - 1, c_int), 1_c_int, opaw_dt, d_dip, both, 2_c_int), &
                      'plot_dip files')
    end if
  end subroutine

  ! drop-in for write_to_restart (src/tddft_write_to_restart.f90:1): the
  ! resident orbitals go to ./tddft_save
  subroutine gpu_write_to_restart(it)
    integer, intent(in) :: it
    call opaw_check(opaw_write_restart(opaw_ctx, './tddft_save'//c_null_char, int(it, c_int)), &
                    'write_to_restart')
  end subroutine

  ! drop-in for read_phi_bar (src/read_phi_bar.f90:1): loads ./tddft_save
  ! into the resident orbitals and returns it_restart
  subroutine gpu_read_phi_bar(it_restart)
    integer, intent(out) :: it_restart
    integer(c_int) :: it
    call opaw_check(opaw_read_restart(opaw_ctx, './tddft_save'//c_null_char, it), 'read_phi_bar')
    it_restart = it
  end subroutine

  ! ground-state SCF drop-ins (src/main.f90:66-95, h_type 1); the orbitals
  ! phit_tot of the single system stay resident between the calls

  ! drop-in for ncpaw_make_hamiltonian (src/ncpaw_make_hamiltonian.f90:1)
  subroutine gpu_ncpaw_make_hamiltonian()
    call opaw_check(opaw_ncpaw_make_hamiltonian(opaw_ctx, 0_c_int, 1_c_int), 'ncpaw_make_hamiltonian')
  end subroutine

  ! drop-in for get_hminmax (src/hminmax.f90:18): niter 25, fctr 1.2
  subroutine gpu_get_hminmax(hmin, hmax)
    real(c_double), intent(out) :: hmin, hmax
    call opaw_check(opaw_hminmax(opaw_ctx, 0_c_int, 1_c_int, 25_c_int, 1.2d0, 12345_c_long_long, hmin, hmax), &
                    'get_hminmax')
  end subroutine

  ! drop-in for chebyr_m(hmax, hbot) (src/chebyr_m.f90:1), nch = 10
  subroutine gpu_chebyr_m(hmax, hbot)
    real(c_double), intent(in) :: hmax, hbot
    call opaw_check(opaw_chebyshev_filter(opaw_ctx, 0_c_int, 1_c_int, 10_c_int, hbot, hmax), 'chebyr_m')
  end subroutine

  ! drop-in for orthog_phi (src/orthog_phi.f90:1)
  subroutine gpu_orthog_phi(ovl)
    real(c_double) :: ovl(*)
    call opaw_check(opaw_orthog_phi(opaw_ctx, 0_c_int, 1_c_int, ovl), 'orthog_phi')
  end subroutine

  ! drop-in for diagh (src/diag.f90:1): eig ascending
  subroutine gpu_diagh(eig)
    real(c_double) :: eig(*)
    call opaw_check(opaw_diagh(opaw_ctx, 0_c_int, 1_c_int, eig), 'diagh')
  end subroutine

end module opaw_hip_mod
