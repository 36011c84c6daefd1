! Drop-ins under the reference's own names and argument lists: a Fortran
! host links this file (with opaw_hip_mod.f90 and libopaw_hip.so) in place of
! src/hphi.f90, src/hphi_pert.f90, src/shs.f90, src/snphi.f90, src/proj.f90,
! src/proj1.f90 and src/rk4.f90.  The fields are the reference's complex*16
! host arrays (nx, ny, nz per orbital, x fastest); each call copies them to
! the device, runs the operator of the context held in opaw_hip_mod (system 0
! for the unperturbed routines, 1 for the _pert ones) and copies the result
! back.  The device keeps vks, D_ij and the projector tables; ik is the
! reference's k index, 1 for the single kinetic table.  Any failure stops
! the run with the library's message, the reference's convention.

! h_phi(pin, hp, ik) (src/hphi.f90:1-22): hp = T pin + vks pin + addvnl
subroutine h_phi(pin, hp, ik)
  use iso_c_binding
  use opaw_hip_mod
  implicit none
  complex(c_double_complex) :: pin(*), hp(*)
  integer :: ik
  call opaw_require_ik(ik, 'h_phi')
  call opaw_check(opaw_h_phi_host(opaw_ctx, 0_c_int, pin, hp, 1_c_int), 'h_phi')
end subroutine

! h_phi_pert(pin, hp, ik) (src/hphi_pert.f90:1): the perturbed system's H
subroutine h_phi_pert(pin, hp, ik)
  use iso_c_binding
  use opaw_hip_mod
  implicit none
  complex(c_double_complex) :: pin(*), hp(*)
  integer :: ik
  call opaw_require_ik(ik, 'h_phi_pert')
  call opaw_check(opaw_h_phi_host(opaw_ctx, 1_c_int, pin, hp, 1_c_int), 'h_phi_pert')
end subroutine

! shs(pin, pout, ik) (src/shs.f90:3-24): pout = S^-1/2 H S^-1/2 pin
subroutine shs(pin, pout, ik)
  use iso_c_binding
  use opaw_hip_mod
  implicit none
  complex(c_double_complex) :: pin(*), pout(*)
  integer :: ik
  call opaw_require_ik(ik, 'shs')
  call opaw_check(opaw_shs_host(opaw_ctx, 0_c_int, pin, pout, 1_c_int), 'shs')
end subroutine

! shs_pert(pin, pout, ik) (src/shs.f90:50-70)
subroutine shs_pert(pin, pout, ik)
  use iso_c_binding
  use opaw_hip_mod
  implicit none
  complex(c_double_complex) :: pin(*), pout(*)
  integer :: ik
  call opaw_require_ik(ik, 'shs_pert')
  call opaw_check(opaw_shs_host(opaw_ctx, 1_c_int, pin, pout, 1_c_int), 'shs_pert')
end subroutine

! sn_phi(pin, sp, ik, n) (src/snphi.f90:1-49): sp = S^n pin
subroutine sn_phi(pin, sp, ik, n)
  use iso_c_binding
  use opaw_hip_mod
  implicit none
  complex(c_double_complex) :: pin(*), sp(*)
  integer :: ik
  real(c_double) :: n
  call opaw_require_ik(ik, 'sn_phi')
  call opaw_check(opaw_sn_phi_host(opaw_ctx, pin, sp, 1_c_int, n), 'sn_phi')
end subroutine

! proj(ia, pin, ca, ms, ik) (src/proj.f90:1-45): ca(is) = dv sum pin conj(p_is)
! over atom ia's box
subroutine proj(ia, pin, ca, ms, ik)
  use iso_c_binding
  use opaw_hip_mod
  implicit none
  integer :: ia, ms, ik
  complex(c_double_complex) :: pin(*), ca(ms)
  call opaw_require_ik(ik, 'proj')
  call opaw_check(opaw_proj_host(opaw_ctx, int(ia, c_int), pin, 0_c_int, ca, int(ms, c_int)), 'proj')
end subroutine

! proj1(ia, pin, ca, ms, ik) (src/proj1.f90:1): the p1 projectors
subroutine proj1(ia, pin, ca, ms, ik)
  use iso_c_binding
  use opaw_hip_mod
  implicit none
  integer :: ia, ms, ik
  complex(c_double_complex) :: pin(*), ca(ms)
  call opaw_require_ik(ik, 'proj1')
  call opaw_check(opaw_proj_host(opaw_ctx, int(ia, c_int), pin, 1_c_int, ca, int(ms, c_int)), 'proj1')
end subroutine

! rk4_prop_shs(p, pp, n, nb) (src/rk4.f90:1-103): one RK4 step of phi_bar (p)
! and phi_bar_pert (pp) with tddft_mod's dt (gpu_set_dt).  p and pp must be
! the orbitals the current Hamiltonian was built from (gpu_scatter_phi_bar,
! then gpu_ncpaw_make_hamiltonian_tddft, as the reference's loop builds H
! from phi_bar before propdt); the library refuses anything else.
subroutine rk4_prop_shs(p, pp, n, nb)
  use iso_c_binding
  use opaw_hip_mod
  implicit none
  integer :: n, nb
  complex(c_double_complex) :: p(n, nb), pp(n, nb)
  call opaw_check(opaw_rk4_prop_shs_host(opaw_ctx, p, pp, int(n, c_long_long), int(nb, c_int)), 'rk4_prop_shs')
end subroutine
