!> md(): one PFMDS run driven through the libpfmds C-ABI.
!>
!> Keeps the reference's interface: the settings-file grammar and group /
!> interaction definitions (md_simulation.f90:48-97, md_interactions.f90:38-136),
!> the extended-xyz input, the integrator schedule with its switching rule and
!> nvms exit (md_simulation.f90:114-147), and the log / out / snapshot /
!> trajectory / final-xyz outputs with their formats
!> (md_simulation.f90:188-271, md_read_write.f90:65-107).  Each integrator's
!> steps run on the GPU in one pfm_run call; the values the outputs need are
!> taken on the device at their steps (pfm_set_records) and written by
!> md_record while later steps run.
module pfm_md
use iso_c_binding
use pfm_binding
implicit none

real(8), parameter :: pfm_mass_coef = 1.6605389217d0/1.6021765654d0*100.d0
real(8), parameter :: pfm_kb = 1.3806488d0/1.6021765654d0*1.d-4

type group_def
  integer :: n
  integer, allocatable :: idx(:)
end type group_def

!> What md_record needs to write md()'s outputs
type md_writer
  integer :: out_id, log_id, out_period, period_traj, traj_g, all_atoms, xyz_moving, ninter, nnhc, n
  integer :: ilog, isnap
  character(len=128) :: prefix
  character(len=32) :: iname, fmt_int, fmt_nhc
  real(8) :: box(3), wall0, wall_out
  type(group_def), allocatable :: groups(:)
  real(8), allocatable :: mass(:), e_inter(:), e_nhc(:)
  character(len=32), allocatable :: types(:)
end type md_writer
type(md_writer), save :: wr

contains

!> One output record (pfm_set_records): the log line, the out block, the
!> snapshot and the trajectory frame due at rec%step, in md()'s formats
!> (md_simulation.f90:188-229, md_read_write.f90:65-107).
integer(c_int) function md_record(user, rec) bind(C)
  type(c_ptr), value :: user
  type(pfm_record), intent(in) :: rec
  real(8), pointer :: p(:), pos(:,:), vel(:,:)
  integer(c_int), pointer :: gn(:)
  type(group_def) :: ga, gt
  real(8) :: e_nhc_sum, total, conserved, fs, mcv, w1, w2
  character(len=128) :: fname
  integer :: t
  call system_clock_seconds(w1)
  t = rec%step
  if (rec%has_energy /= 0) then
    if (wr%ninter > 0) then
      call c_f_pointer(rec%e_inter, p, [wr%ninter])
      wr%e_inter(1:wr%ninter) = p
    endif
    e_nhc_sum = 0.d0
    if (wr%nnhc > 0) then
      call c_f_pointer(rec%e_nhc, p, [wr%nnhc])
      wr%e_nhc(1:wr%nnhc) = p
      e_nhc_sum = sum(wr%e_nhc(1:wr%nnhc))
    endif
    total = rec%pe + rec%ke
    conserved = total + e_nhc_sum
    if (mod(t, wr%ilog) == 0) then
      write(wr%log_id,'(A6,i9,7f24.6)',advance='no') trim(wr%iname), t, rec%sim_time, conserved, e_nhc_sum, &
          total, rec%pe, rec%ke, rec%temp
      write(wr%log_id,wr%fmt_int,advance='no') wr%e_inter(1:wr%ninter)
      write(wr%log_id,wr%fmt_nhc) wr%e_nhc(1:wr%nnhc)
    endif
    if (mod(t, wr%out_period) == 0) then
      fs = sqrt(sum(rec%force_sum**2))
      mcv = sqrt(sum(rec%mass_center_vel**2))
      write(wr%out_id,'(A,f10.2)') 'exe time (s) = ', w1 - wr%wall0
      write(wr%out_id,'(A,i12,A,A6)') 'step = ', t, ' integrator = ', trim(wr%iname)
      write(wr%out_id,'(A,es21.9)') 'conserved energy (eV) = ', conserved
      write(wr%out_id,'(A,es16.6)') 'forces sum (eV/A) = ', fs
      write(wr%out_id,'(A,es16.6)') 'c.o.m. velocity (A/fs) = ', mcv
      write(wr%out_id,'(A,3f12.6)') 'c.o.m. position (A) = ', rec%mass_center
      write(wr%out_id,'(A,f12.6)') 'maximum velocity (A/fs) = ', rec%max_vel
      if (trim(wr%iname) /= 'nvms' .and. fs > 1.d-14) write(wr%out_id,*) 'WARNING: forces sum is too big'
      if (trim(wr%iname) == 'nvt' .and. wr%xyz_moving == wr%all_atoms .and. mcv > 1.d-14) &
          write(wr%out_id,*) 'WARNING: center of mass velocity is too big'
      write(wr%out_id,*)
    endif
  endif
  if (rec%has_frame /= 0) then
    call c_f_pointer(rec%pos, pos, [3, wr%n])
    call c_f_pointer(rec%vel, vel, [3, wr%n])
    ! the groups' active atoms at the step (change_particle_group_N)
    call c_f_pointer(rec%group_n, gn, [33])
    ga = wr%groups(wr%all_atoms)
    ga%n = gn(wr%all_atoms + 1)
    gt = wr%groups(wr%traj_g)
    gt%n = gn(wr%traj_g + 1)
    if (mod(t, wr%isnap) == 0 .and. t /= 0) then
      write(fname,'(A,A,i6.6,A)') trim(wr%prefix), 'snapshot_', t, '.xyz'
      call write_snapshot(trim(fname), ga, wr%box, pos, vel, wr%mass, wr%types)
    endif
    if (mod(t, wr%period_traj) == 0) then
      write(fname,'(A,A,i2.2,A)') trim(wr%prefix), 'traj_', wr%traj_g, '.xyz'
      call append_trajectory(trim(fname), gt, wr%box, pos, vel, wr%mass, wr%types, t)
    endif
  endif
  call system_clock_seconds(w2)
  wr%wall_out = wr%wall_out + (w2 - w1)
  md_record = 0
end function md_record

integer function next_multiple(step, period)
  integer :: step, period
  next_multiple = ((step + period - 1)/period)*period
end function next_multiple

!> Atoms whose species is one of `names`, by name then ascending index
!> (create_particle_group, md_general.f90:57-80).
subroutine make_group(g, names, types)
  type(group_def) :: g
  character(len=32) :: names(:), types(:)
  integer :: j, i, k
  g%n = 0
  do j = 1, size(names)
    g%n = g%n + count(types == names(j))
  enddo
  allocate(g%idx(max(g%n, 1)))
  k = 0
  do j = 1, size(names)
    do i = 1, size(types)
      if (types(i) == names(j)) then
        k = k + 1
        g%idx(k) = i
      endif
    enddo
  enddo
end subroutine make_group

subroutine read_xyz(path, box, pos, vel, mass, types)
  character(len=*) :: path
  real(8) :: box(3), bm(9)
  real(8), allocatable :: pos(:,:), vel(:,:), mass(:)
  character(len=32), allocatable :: types(:)
  character(len=32) :: c
  integer :: n, i, u
  open(newunit=u, file=path, status='old')
  read(u,*) n
  read(u,*) c, bm
  box = (/ bm(1), bm(5), bm(9) /)
  allocate(pos(3,n), vel(3,n), mass(n), types(n))
  do i = 1, n
    read(u,*) pos(:,i), vel(:,i), mass(i), types(i)
  enddo
  close(u)
end subroutine read_xyz

!> write_particle_group / write_particle_group_append formats (md_read_write.f90:65-107)
subroutine write_snapshot(path, g, box, pos, vel, mass, types)
  character(len=*) :: path
  type(group_def) :: g
  real(8) :: box(3), pos(:,:), vel(:,:), mass(:)
  character(len=32) :: types(:)
  integer :: u, k, i
  open(newunit=u, file=path)
  write(u,*) g%n
  write(u,'(A,9f16.6,A)') 'Lattice="', box(1), 0., 0., 0., box(2), 0., 0., 0., box(3), &
      ' " Properties=pos:R:3:vel:R:3:mass:R:1:species:S:1'
  do k = 1, g%n
    i = g%idx(k)
    write(u,'(7f27.16,A,A)') pos(:,i), vel(:,i), mass(i), '    ', types(i)
  enddo
  close(u)
end subroutine write_snapshot

subroutine append_trajectory(path, g, box, pos, vel, mass, types, step)
  character(len=*) :: path
  type(group_def) :: g
  real(8) :: box(3), pos(:,:), vel(:,:), mass(:)
  character(len=32) :: types(:)
  integer :: u, k, i, step
  if (g%n <= 0) return
  open(newunit=u, file=path, action='write', position='append')
  write(u,*) g%n
  write(u,'(A,i9,A,9f10.4,A)') 'time_step: ', step, '    Lattice="', box(1), 0., 0., 0., box(2), 0., 0., 0., &
      box(3), ' " Properties=pos:R:3:vel:R:3:mass:R:1:species:S:1'
  do k = 1, g%n
    i = g%idx(k)
    write(u,'(7f10.4,A,A)') pos(:,i), vel(:,i), mass(i), '    ', trim(types(i))
  enddo
  close(u)
end subroutine append_trajectory

!> Parameters of one interaction in read_*_parameters order.
subroutine read_params(path, kind, p, np)
  character(len=*) :: path
  integer :: kind, np, u
  real(8) :: p(24)
  logical :: simp
  p = 0.d0
  open(newunit=u, file=path, status='old')
  select case (kind)
  case (PFM_LJ, PFM_LJ1G)
    read(u,*) p(1), p(2)
    read(u,*) p(3), p(4)
    np = 4
  case (PFM_LJC)
    read(u,*) p(1), p(2), p(3)
    read(u,*) p(4), p(5)
    read(u,*) simp
    p(6) = merge(1.d0, 0.d0, simp)
    np = 6
  case (PFM_MORSEC)
    read(u,*) p(1), p(2), p(3), p(4)
    read(u,*) p(5), p(6)
    read(u,*) simp
    p(7) = merge(1.d0, 0.d0, simp)
    np = 7
  case (PFM_RJL)
    read(u,*) p(1), p(2), p(3), p(4), p(5)
    read(u,*) p(6), p(7)
    np = 7
  case (PFM_TB)
    read(u,*) p(1), p(2), p(3), p(4), p(5), p(6), p(7), p(8)
    read(u,*) p(9), p(10)
    np = 10
  case (PFM_REBOSC)
    ! read_REBOsc_parameters (REBOsolidcarbon.f90:17-22)
    read(u,*) p(1), p(2), p(3)
    read(u,*) p(4), p(5), p(6)
    read(u,*) p(7), p(8), p(9)
    read(u,*) p(10)
    read(u,*) p(11), p(12), p(13), p(14), p(15), p(16)
    read(u,*) p(17), p(18)
    np = 18
  end select
  close(u)
end subroutine read_params

!> Seeded Maxwell velocities for new_velocities=T, scaled to `temp` over the
!> group with 3N degrees of freedom after removing the centre-of-mass
!> velocity (set_new_temperature, md_general.f90:328-340; the generator is
!> the compiler's random_number, seeded from rand_seed).
subroutine new_velocities(vel, mass, g, temp, rand_seed)
  real(8) :: vel(:,:), mass(:), temp, a1, a2, b, mcv(3), tm, ke, t
  type(group_def) :: g
  integer :: rand_seed, k, i, c, ns
  integer, allocatable :: seed(:)
  call random_seed(size=ns)
  allocate(seed(ns))
  do k = 1, ns
    seed(k) = rand_seed*7919 + 104729*k
  enddo
  call random_seed(put=seed)
  do k = 1, g%n
    i = g%idx(k)
    do c = 1, 3
      b = 2.d0
      do while (b >= 1.d0 .or. b == 0.d0)
        call random_number(a1)
        call random_number(a2)
        a1 = 2.d0*a1 - 1.d0
        a2 = 2.d0*a2 - 1.d0
        b = a1**2 + a2**2
      enddo
      vel(c,i) = a1*sqrt(-2.d0*log(b)/b)*sqrt(1.3806488d0/1.6605389217d0*1.d-6/mass(i))
    enddo
  enddo
  mcv = 0.d0
  tm = 0.d0
  do k = 1, g%n
    i = g%idx(k)
    mcv = mcv + mass(i)*vel(:,i)
    tm = tm + mass(i)
  enddo
  mcv = mcv/tm
  ke = 0.d0
  do k = 1, g%n
    i = g%idx(k)
    vel(:,i) = vel(:,i) - mcv
    ke = ke + mass(i)*sum(vel(:,i)**2)/2*pfm_mass_coef
  enddo
  t = 2*ke/pfm_kb/(3*g%n)
  if (t > 0.d0) then
    do k = 1, g%n
      i = g%idx(k)
      vel(:,i) = vel(:,i)*sqrt(temp/t)
    enddo
  endif
end subroutine new_velocities

subroutine md(out_id, all_out_id, input_path, settings_filename, output_prefix, out_period, num_of_omp_treads, &
              rand_seed, precision)
  integer :: out_id, all_out_id, out_period, num_of_omp_treads, rand_seed, precision
  character(len=*) :: input_path, settings_filename, output_prefix
  type(c_ptr) :: h
  integer(c_int) :: rc, last, ev, nact
  integer :: fid, log_id, i, j, k, n, ntypes, ngroups, nint, nnhc, ninter, nchange, md_step, md_step_limit
  integer :: zmp, all_moving, xyz_moving, z_moving, all_atoms, traj_g, period_traj, idx_int, cend, np, nl_n
  integer :: nhc_g, nhc_m, kind, integ_kind
  integer, allocatable :: ilen(:), isnap(:), ilog(:), chg(:,:)
  real(8), allocatable :: idt(:)
  character(len=32), allocatable :: iname(:), names(:,:), types(:)
  real(8), allocatable :: pos(:,:), vel(:,:), frc(:,:), mass(:), e_inter(:), e_nhc(:)
  real(8) :: box(3), ms_de, t0, nhc_t, nhc_q1, p(24), rc_(3), pe, ke, temp, sim_time, total
  real(8) :: wall0, wall_md, wall_out, wtmp, dpe
  integer(c_int) :: gn(6), nmax(3), per(3), eper(2), xper(2)
  logical :: newv, invert
  character(len=128) :: str, logfilename, init_xyz, pfile, fname
  character(len=32) :: iname_cur, fmt_int, fmt_nhc, inter_name
  type(group_def), allocatable :: groups(:)

  call system_clock_seconds(wall0)
  open(newunit=fid, file=trim(input_path)//trim(settings_filename), status='old')
  write(out_id,'(A,A)') 'settings_filename: ', trim(settings_filename)
  read(fid,*) str, md_step_limit;  write(out_id,'(A32,i12)') str, md_step_limit
  read(fid,*) str, logfilename;    write(out_id,'(A32,A,A)') str, '	', trim(logfilename)
  read(fid,*) str, init_xyz;       write(out_id,'(A32,A,A)') str, '	', trim(init_xyz)
  call read_xyz(trim(input_path)//trim(init_xyz), box, pos, vel, mass, types)
  n = size(mass)
  allocate(frc(3,n))
  frc = 0.d0
  write(out_id,'(A,3f16.6)') 'box_size: ', box
  write(out_id,'(A,i12)') 'particles_num: ', n
  read(fid,*) str, newv;           write(out_id,'(A32,l8)') str, newv
  read(fid,*) str, zmp;            write(out_id,'(A32,i12)') str, zmp
  read(fid,*) str, ntypes;         write(out_id,'(A32,i12)') str, ntypes
  read(fid,*) str, ngroups;        write(out_id,'(A32,i12)') str, ngroups
  allocate(names(ntypes, ngroups), groups(ngroups))
  do i = 1, ngroups
    read(fid,*) str, names(:,i)
    call make_group(groups(i), names(:,i), types)
    write(out_id,'(i6,A,*(A12))', advance='no') i, ' ', names(:,i)
    write(out_id,'(i9)') groups(i)%n
  enddo
  read(fid,*) str, all_moving;     write(out_id,'(A32,i12)') str, all_moving
  read(fid,*) str, xyz_moving;     write(out_id,'(A32,i12)') str, xyz_moving
  read(fid,*) str, z_moving;       write(out_id,'(A32,i12)') str, z_moving
  read(fid,*) str, all_atoms;      write(out_id,'(A32,i12)') str, all_atoms
  read(fid,*) str, traj_g;         write(out_id,'(A32,i12)') str, traj_g
  read(fid,*) str, period_traj;    write(out_id,'(A32,i12)') str, period_traj
  read(fid,*) str, nchange;        write(out_id,'(A32,i12)') str, nchange
  ! runtime group changes (md_simulation.f90:63-70)
  allocate(chg(5, max(nchange, 1)))
  do i = 1, nchange
    read(fid,*) str, chg(1,i), chg(2,i);          write(out_id,'(A32,2i12)') str, chg(1,i), chg(2,i)
    read(fid,*) str, chg(3,i), chg(4,i), chg(5,i); write(out_id,'(A32,3i12)') str, chg(3,i), chg(4,i), chg(5,i)
  enddo
  read(fid,*) str, invert;         write(out_id,'(A32,l8)') str, invert
  read(fid,*) str, nint;           write(out_id,'(A32,i12)') str, nint
  allocate(iname(0:nint), idt(0:nint), ilen(0:nint), isnap(0:nint), ilog(0:nint))
  iname(0) = 'none'; idt(0) = 0.d0; ilen(0) = 0; isnap(0) = 1; ilog(0) = 1
  read(fid,'(A)') str;             write(out_id,'(A)') trim(str)
  do i = 1, nint
    read(fid,*) iname(i), idt(i), ilen(i), isnap(i), ilog(i)
    write(out_id,'(A,A6,f10.5,4i9)') '  ', iname(i), idt(i), ilen(i), isnap(i), ilog(i)
  enddo
  read(fid,*) str, ms_de;          write(out_id,'(A32,es16.6)') str, ms_de
  read(fid,*) str, nnhc;           write(out_id,'(A32,i12)') str, nnhc
  rc = pfm_create(h, int(n, c_int), box, pos, vel, mass, int(precision, c_int))
  call pfm_check(h, rc, out_id)
  ! groups are known once the xyz file and the group table are read
  do i = 1, ngroups
    if (groups(i)%n > 0) then
      rc = pfm_set_group(h, int(i, c_int), int(groups(i)%n, c_int), groups(i)%idx)
    else
      rc = pfm_set_group(h, int(i, c_int), 0_c_int, groups(i)%idx)
    endif
    call pfm_check(h, rc, out_id)
  enddo
  do i = 1, nchange
    rc = pfm_add_group_change(h, int(chg(1,i), c_int), int(chg(2,i), c_int), int(chg(3,i), c_int), &
                              int(chg(4,i), c_int), int(chg(5,i), c_int))
    call pfm_check(h, rc, out_id)
  enddo
  do i = 1, nnhc
    read(fid,*) nhc_g, nhc_t, nhc_m, nhc_q1
    write(out_id,'(i6,f16.6,i6,f16.6)') nhc_g, nhc_t, nhc_m, nhc_q1
    rc = pfm_add_nhc(h, int(nhc_g, c_int), nhc_t, int(nhc_m, c_int), nhc_q1)
    call pfm_check(h, rc, out_id)
  enddo
  read(fid,*) str, t0;             write(out_id,'(A32,f16.6)') str, t0
  if (t0 < 0.d0) t0 = 0.d0
  if (newv) then
    call new_velocities(vel, mass, groups(all_moving), t0, rand_seed)
    rc = pfm_set_state_vel(h, pos, vel)
    call pfm_check(h, rc, out_id)
  endif
  rc = pfm_set_roles(h, int(all_moving, c_int), int(xyz_moving, c_int), int(z_moving, c_int), &
                     int(all_atoms, c_int), int(zmp, c_int), merge(1_c_int, 0_c_int, invert))
  call pfm_check(h, rc, out_id)
  read(fid,*) str, ninter;         write(out_id,'(A32,i12)') str, ninter
  do i = 1, ninter
    read(fid,*) inter_name, pfile
    select case (trim(inter_name))
    case ('lj');     kind = PFM_LJ;     nl_n = 2
    case ('lj1g');   kind = PFM_LJ1G;   nl_n = 1
    case ('ljc');    kind = PFM_LJC;    nl_n = 3
    case ('morsec'); kind = PFM_MORSEC; nl_n = 3
    case ('rjl');    kind = PFM_RJL;    nl_n = 1
    case ('tb');     kind = PFM_TB;     nl_n = 1
    case ('rebosc'); kind = PFM_REBOSC; nl_n = 1
    case default
      write(out_id,*) 'error: unknown interaction name ', trim(inter_name)
      stop 1
    end select
    write(out_id,'(2A32,i6)') inter_name, pfile, nl_n
    call read_params(trim(input_path)//trim(pfile), kind, p, np)
    do j = 1, nl_n
      read(fid,*) gn(2*j-1), gn(2*j), nmax(j), rc_(j), per(j)
      write(out_id,'(3i6,f16.6,i9)') gn(2*j-1), gn(2*j), nmax(j), rc_(j), per(j)
    enddo
    rc = pfm_add_interaction(h, int(kind, c_int), p, int(np, c_int), int(nl_n, c_int), gn, nmax, rc_, per)
    call pfm_check(h, rc, out_id)
  enddo
  close(fid)
  allocate(e_inter(max(ninter,1)), e_nhc(max(nnhc,1)))
  e_inter = 0.d0
  e_nhc = 0.d0
  ! energy columns of the log and summary lines: any count, f20.9 each
  fmt_int = '(*(f20.9))'
  fmt_nhc = '(*(f20.9))'
  open(newunit=log_id, file=trim(output_prefix)//trim(logfilename))
  write(out_id,'(A24,i6,A)') 'RUNNING ON ', num_of_omp_treads, ' OPENMP THREADS'
  write(out_id,*)

  wr%out_id = out_id
  wr%log_id = log_id
  wr%out_period = out_period
  wr%period_traj = period_traj
  wr%traj_g = traj_g
  wr%all_atoms = all_atoms
  wr%xyz_moving = xyz_moving
  wr%ninter = ninter
  wr%nnhc = nnhc
  wr%n = n
  wr%prefix = output_prefix
  wr%fmt_int = fmt_int
  wr%fmt_nhc = fmt_nhc
  wr%box = box
  wr%wall0 = wall0
  wr%wall_out = 0.d0
  wr%groups = groups
  wr%mass = mass
  wr%types = types
  allocate(wr%e_inter(max(ninter,1)), wr%e_nhc(max(nnhc,1)))
  wr%e_inter = 0.d0
  wr%e_nhc = 0.d0

  call system_clock_seconds(wall_md)
  md_step = 0
  idx_int = 0
  iname_cur = 'none'
  integ_kind = PFM_NVE
  do while (md_step <= md_step_limit)
    ! integrator switch (md_simulation.f90:121-136)
    if (idx_int == 0 .or. md_step - 1 == sum(ilen(0:idx_int))) then
      do i = idx_int + 1, nint
        if (ilen(i) > 0) exit
      enddo
      if (i > nint) then
        write(out_id,*) 'no more integrators'
        exit
      endif
      idx_int = i
      iname_cur = iname(i)
      select case (trim(iname_cur))
      case ('nvt');  integ_kind = PFM_NVT
      case ('nve');  integ_kind = PFM_NVE
      case ('nvms'); integ_kind = PFM_NVMS
      case default
        write(out_id,*) 'error: unknown integrator ', trim(iname_cur)
        stop 1
      end select
    endif
    cend = max(min(sum(ilen(0:idx_int)), md_step_limit), md_step)
    ! log and out steps take energy records, snapshot and trajectory steps frames
    wr%iname = iname_cur
    wr%ilog = ilog(idx_int)
    wr%isnap = isnap(idx_int)
    eper = (/ int(ilog(idx_int), c_int), int(out_period, c_int) /)
    xper = (/ int(isnap(idx_int), c_int), int(period_traj, c_int) /)
    rc = pfm_set_records(h, 2_c_int, eper, 2_c_int, xper, 4_c_int, c_funloc(md_record), c_null_ptr)
    call pfm_check(h, rc, out_id)
    rc = pfm_run(h, int(integ_kind, c_int), idt(idx_int), int(md_step, c_int), int(cend, c_int), 0_c_int, &
                 ms_de, last, ev)
    call pfm_check(h, rc, out_id)
    if (ev == PFM_EV_CONVERGED) then
      write(out_id,*) 'potential energy diffrence is small enough'
      md_step = last
      exit
    endif
    md_step = cend + 1
  enddo
  wall_out = wr%wall_out
  call system_clock_seconds(wtmp)
  wall_md = wtmp - wall_md

  if (trim(iname_cur) == 'nvms') then
    rc = pfm_pe_change(h, dpe)
    write(out_id,*) 'potential energy difference: ', dpe
  endif
  write(out_id,*) 'steps number:', md_step - 1
  write(out_id,*)
  write(out_id,*) 'PERFOMANCE:'
  write(out_id,'(A24,f10.2,A,f10.2,A)') 'MD:', wall_md, ' S ', 100.d0, '%'
  write(out_id,'(A24,f10.2,A,f10.2,A)') 'GPU STEPS:', wall_md - wall_out, ' S ', &
      (wall_md - wall_out)/max(wall_md, 1.d-30)*100, '%'
  write(out_id,'(A24,f10.2,A,f10.2,A)') 'OUTPUT:', wall_out, ' S ', wall_out/max(wall_md, 1.d-30)*100, '%'
  write(out_id,'(A24,f16.2)') 'TIME STEPS PER HOUR:', md_step/max(wall_md, 1.d-30)*3600
  write(out_id,*)

  rc = pfm_energies(h, e_inter, pe, ke, temp, e_nhc, sim_time)
  total = pe + ke
  write(all_out_id,'(A32,i9,5f20.9)',advance='no') trim(output_prefix), md_step - 1, sim_time, total, pe, ke, temp
  write(all_out_id,fmt_int,advance='no') e_inter(1:ninter)
  write(log_id,'(A32,i9,5f20.9)',advance='no') trim(output_prefix), md_step - 1, sim_time, total, pe, ke, temp
  write(log_id,fmt_int,advance='no') e_inter(1:ninter)
  if (md_step /= 0) then
    rc = pfm_get_state(h, pos, vel, frc)
    rc = pfm_group_count(h, int(all_atoms, c_int), nact)
    groups(all_atoms)%n = nact
    write(fname,'(A,A,A)') trim(output_prefix), 'final_', trim(init_xyz)
    call write_snapshot(trim(fname), groups(all_atoms), box, pos, vel, mass, types)
  endif
  close(log_id)
  rc = pfm_destroy(h)
end subroutine md

subroutine system_clock_seconds(s)
  real(8) :: s
  integer(8) :: c, r
  call system_clock(c, r)
  s = real(c, 8)/real(r, 8)
end subroutine system_clock_seconds

integer(c_int) function pfm_set_state_vel(h, pos, vel)
  type(c_ptr) :: h
  real(8) :: pos(:,:), vel(:,:)
  interface
    integer(c_int) function c_set_state(hh, p, v) bind(C, name='pfm_set_state')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: hh
      real(c_double) :: p(*), v(*)
    end function
  end interface
  pfm_set_state_vel = c_set_state(h, pos, vel)
end function pfm_set_state_vel

end module pfm_md
