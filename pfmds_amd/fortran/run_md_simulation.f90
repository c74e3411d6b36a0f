!> Command-line front end: the flags of the reference runner
!> (run_md_simulation.f90:26-76: -i -p -op -omp_n -ipath -ilist -opath -ofile)
!> plus -prec fp64|mixed for the precision of the GPU force arithmetic.
program md_run
use pfm_binding, only: PFM_FP64, PFM_MIXED
use pfm_md
implicit none
integer :: i, out_period, nthreads, out_id, all_out_id, set_num, precision
character(len=128) :: settings_filename, settings_files_list, all_out_file, output_prefix, input_path, out_path, str
character(len=32) :: arg
character(len=80) :: line

line = repeat('_', 80)
out_id = 6
all_out_id = 66
out_period = 1
settings_files_list = ''
settings_filename = 'md_run_settings.txt'
all_out_file = 'all_out.txt'
output_prefix = ''
input_path = ''
out_path = ''
nthreads = 1
precision = PFM_FP64
write(out_id,'(A)') line
i = 0
do while (i < command_argument_count())
  i = i + 1
  call get_command_argument(i, arg)
  select case (arg)
  case ('-i', '--input')
    i = i + 1; call get_command_argument(i, settings_filename)
  case ('-p', '--prefix')
    i = i + 1; call get_command_argument(i, output_prefix)
    write(out_id,'(A,A)') 'output_prefix:  ', trim(output_prefix)
  case ('-op', '--out_period')
    i = i + 1; call get_command_argument(i, str)
    write(out_id,'(A,A)') 'out_period:         ', trim(str)
    read(str,*) out_period
  case ('-omp_n', '--openmp_threads_num')
    i = i + 1; call get_command_argument(i, str)
    write(out_id,'(A,A)') 'openmp_threads_num: ', trim(str)
    read(str,*) nthreads
  case ('-ipath', '--input_path')
    i = i + 1; call get_command_argument(i, input_path)
    write(out_id,'(A,A)') 'input_path:         ', trim(input_path)
  case ('-ilist', '--input_list')
    i = i + 1; call get_command_argument(i, settings_files_list)
    write(out_id,'(A,A)') 'input_list:         ', trim(settings_files_list)
  case ('-opath', '--out_path')
    i = i + 1; call get_command_argument(i, out_path)
    write(out_id,'(A,A)') 'out_path:           ', trim(out_path)
  case ('-ofile', '--all_out_file')
    i = i + 1; call get_command_argument(i, all_out_file)
    write(out_id,'(A,A)') 'all_out_file:   ', trim(all_out_file)
  case ('-prec', '--precision')
    i = i + 1; call get_command_argument(i, str)
    write(out_id,'(A,A)') 'precision:          ', trim(str)
    precision = PFM_FP64
    if (trim(str) == 'mixed') precision = PFM_MIXED
  end select
enddo
if (settings_files_list /= '') then
  open(1234, file=trim(input_path)//trim(settings_files_list), status='old')
  open(all_out_id, file=trim(out_path)//trim(all_out_file))
  read(1234,*) set_num
  do i = 1, set_num
    read(1234,*) settings_filename, str
    write(output_prefix,'(A,A)') trim(out_path), trim(str)
    write(out_id,'(A)') line
    write(out_id,'(i6,A,A32,A32)') i, '	', settings_filename, output_prefix
    call md(out_id, all_out_id, input_path, settings_filename, output_prefix, out_period, nthreads, 1, precision)
    write(all_out_id,*)
    write(out_id,'(A)') line
  enddo
  close(1234)
  close(all_out_id)
else
  write(out_id,'(A)') line
  call md(out_id, out_id, input_path, settings_filename, output_prefix, out_period, nthreads, 1, precision)
  write(out_id,*)
  write(out_id,'(A)') line
endif
write(out_id,'(A)') line
end program md_run
