!> ISO_C_BINDING interfaces of the libpfmds C-ABI (include/pfm.h).
!> Positions, velocities and forces cross the boundary as the (3,N) REAL(8)
!> arrays md() already keeps (C double[N][3]); indices are 1-based.
module pfm_binding
use iso_c_binding
implicit none

integer(c_int), parameter :: PFM_OK = 0
integer(c_int), parameter :: PFM_LJ = 1, PFM_LJ1G = 2, PFM_LJC = 3, PFM_MORSEC = 4, PFM_RJL = 5, PFM_TB = 6, &
                              PFM_REBOSC = 8
integer(c_int), parameter :: PFM_NVT = 1, PFM_NVE = 2, PFM_NVMS = 3
integer(c_int), parameter :: PFM_FP64 = 0, PFM_MIXED = 1
integer(c_int), parameter :: PFM_EV_DONE = 0, PFM_EV_CONVERGED = 1

!> struct pfm_record: one output record handed to the pfm_set_records callback
type, bind(C) :: pfm_record
  integer(c_int) :: step, integrator, has_energy, has_frame
  real(c_double) :: sim_time, pe, ke, temp
  real(c_double) :: force_sum(3), mass_center(3), mass_center_vel(3)
  real(c_double) :: max_vel
  type(c_ptr) :: e_inter, e_nhc, pos, vel
  type(c_ptr) :: group_n
end type pfm_record

interface
  integer(c_int) function pfm_create(h, n, box, pos, vel, mass, precision) bind(C, name='pfm_create')
    import :: c_ptr, c_int, c_double
    type(c_ptr) :: h
    integer(c_int), value :: n, precision
    real(c_double) :: box(3), pos(*), vel(*), mass(*)
  end function
  integer(c_int) function pfm_set_group(h, gid, n, idx) bind(C, name='pfm_set_group')
    import :: c_ptr, c_int
    type(c_ptr), value :: h
    integer(c_int), value :: gid, n
    integer(c_int) :: idx(*)
  end function
  integer(c_int) function pfm_set_roles(h, all_moving, xyz_moving, z_moving, all_atoms, zmp, invert_z) &
      bind(C, name='pfm_set_roles')
    import :: c_ptr, c_int
    type(c_ptr), value :: h
    integer(c_int), value :: all_moving, xyz_moving, z_moving, all_atoms, zmp, invert_z
  end function
  integer(c_int) function pfm_add_interaction(h, kind, params, np, nl_n, groups, nmax, rcut, period) &
      bind(C, name='pfm_add_interaction')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    integer(c_int), value :: kind, np, nl_n
    real(c_double) :: params(*), rcut(*)
    integer(c_int) :: groups(*), nmax(*), period(*)
  end function
  integer(c_int) function pfm_add_nhc(h, group, temperature, m, q1) bind(C, name='pfm_add_nhc')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    integer(c_int), value :: group, m
    real(c_double), value :: temperature, q1
  end function
  integer(c_int) function pfm_run(h, integrator, dt, step_begin, step_end, energy_at_end, ms_de, last_step, event) &
      bind(C, name='pfm_run')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    integer(c_int), value :: integrator, step_begin, step_end, energy_at_end
    real(c_double), value :: dt, ms_de
    integer(c_int) :: last_step, event
  end function
  integer(c_int) function pfm_energies(h, e_inter, pe, ke, temp, e_nhc, sim_time) bind(C, name='pfm_energies')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    real(c_double) :: e_inter(*), pe, ke, temp, e_nhc(*), sim_time
  end function
  integer(c_int) function pfm_get_state(h, pos, vel, force) bind(C, name='pfm_get_state')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    real(c_double) :: pos(*), vel(*), force(*)
  end function
  integer(c_int) function pfm_last_error(h, buf, length) bind(C, name='pfm_last_error')
    import :: c_ptr, c_int, c_char
    type(c_ptr), value :: h
    character(kind=c_char) :: buf(*)
    integer(c_int), value :: length
  end function
  integer(c_int) function pfm_set_records(h, n_e, e_periods, n_x, x_periods, frame_slots, fn, user) &
      bind(C, name='pfm_set_records')
    import :: c_ptr, c_int, c_funptr
    type(c_ptr), value :: h, user
    integer(c_int), value :: n_e, n_x, frame_slots
    integer(c_int) :: e_periods(*), x_periods(*)
    type(c_funptr), value :: fn
  end function
  integer(c_int) function pfm_add_group_change(h, gfrom, gto, ts1, ts2, frec) bind(C, name='pfm_add_group_change')
    import :: c_ptr, c_int
    type(c_ptr), value :: h
    integer(c_int), value :: gfrom, gto, ts1, ts2, frec
  end function
  integer(c_int) function pfm_group_count(h, gid, n) bind(C, name='pfm_group_count')
    import :: c_ptr, c_int
    type(c_ptr), value :: h
    integer(c_int), value :: gid
    integer(c_int) :: n
  end function
  integer(c_int) function pfm_pe_change(h, de) bind(C, name='pfm_pe_change')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    real(c_double) :: de
  end function
  integer(c_int) function pfm_destroy(h) bind(C, name='pfm_destroy')
    import :: c_ptr, c_int
    type(c_ptr), value :: h
  end function
end interface

contains

!> Print the library's message for a non-zero status and stop, as the
!> reference does for its fatal conditions.
subroutine pfm_check(h, rc, out_id)
  type(c_ptr) :: h
  integer(c_int) :: rc
  integer :: out_id, k
  character(kind=c_char) :: buf(512)
  character(len=512) :: msg
  if (rc == PFM_OK) return
  k = pfm_last_error(h, buf, 512_c_int)
  msg = ''
  do k = 1, 512
    if (buf(k) == c_null_char) exit
    msg(k:k) = buf(k)
  enddo
  write(out_id, '(A)') trim(msg)
  write(*, '(A)') trim(msg)
  stop 1
end subroutine pfm_check

end module pfm_binding
