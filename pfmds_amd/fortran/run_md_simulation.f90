!> run_md_simulation -- command-line driver of the GPU molecular-dynamics
!> front end.
!>
!> Kept interface (runners/run_md_simulation.f90:26-76): the flags
!>   -i/--input -p/--prefix -op/--out_period -omp_n/--openmp_threads_num
!>   -ipath/--input_path -ilist/--input_list -opath/--out_path -ofile/--all_out_file,
!> the echo line printed for each of them, the separator rules printed around
!> every md() run, the (count, then "settings prefix" rows) list file, and
!> random seed 1.  Added: -prec/--precision fp64|mixed, the arithmetic of the
!> GPU force passes.
!>
!> The options are read through one table (flag spellings, echo label and its
!> width), and a single job loop serves both modes: one settings file, whose
!> closing blank line goes to standard output, or a list of them, each job
!> numbered on standard output with its closing blank line in the summary file.
module md_cli
  use omp_lib, only: omp_get_max_threads
  use pfm_binding, only: PFM_FP64, PFM_MIXED
  implicit none
  private
  public :: cli_options, md_job, read_cli, plan_jobs, rule
  public :: K_IPATH, K_ILIST, K_OPATH, K_OFILE

  integer, parameter :: NOPT = 9
  ! option keys, in table order
  integer, parameter :: K_IN = 1, K_PREFIX = 2, K_OUTP = 3, K_OMP = 4, K_IPATH = 5, K_ILIST = 6, &
                        K_OPATH = 7, K_OFILE = 8, K_PREC = 9
  character(len=8), parameter :: short_flag(NOPT) = [character(len=8) :: &
    '-i', '-p', '-op', '-omp_n', '-ipath', '-ilist', '-opath', '-ofile', '-prec']
  character(len=24), parameter :: long_flag(NOPT) = [character(len=24) :: &
    '--input', '--prefix', '--out_period', '--openmp_threads_num', '--input_path', '--input_list', &
    '--out_path', '--all_out_file', '--precision']
  ! echo label and the column its value starts in (0: the option is not echoed)
  character(len=20), parameter :: echo_label(NOPT) = [character(len=20) :: &
    '', 'output_prefix:', 'out_period:', 'openmp_threads_num:', 'input_path:', 'input_list:', &
    'out_path:', 'all_out_file:', 'precision:']
  integer, parameter :: echo_width(NOPT) = [0, 16, 20, 20, 20, 20, 20, 16, 20]

  type cli_options
    character(len=128) :: value(NOPT)
    integer :: out_period = 1
    integer :: threads = 1
    integer :: precision = PFM_FP64
  end type cli_options

  type md_job
    character(len=128) :: settings = ''
    character(len=128) :: prefix = ''
  end type md_job

contains

  !> the 80-character separator rule
  function rule() result(r)
    character(len=80) :: r
    r = repeat('_', 80)
  end function rule

  integer function option_key(flag)
    character(len=*), intent(in) :: flag
    integer :: k
    option_key = 0
    do k = 1, NOPT
      if (flag == short_flag(k) .or. flag == long_flag(k)) then
        option_key = k
        return
      endif
    enddo
  end function option_key

  !> Reads the command line into o, echoing each recognised option on unit
  !> `out` as it is met; unknown words are skipped.
  subroutine read_cli(o, out)
    type(cli_options), intent(out) :: o
    integer, intent(in) :: out
    integer :: pos, nargs, key
    character(len=128) :: word

    o%value = ''
    o%value(K_IN) = 'md_run_settings.txt'
    o%value(K_OFILE) = 'all_out.txt'
    o%value(K_PREC) = 'fp64'
    o%threads = omp_get_max_threads()
    nargs = command_argument_count()
    pos = 1
    do while (pos <= nargs)
      call get_command_argument(pos, word)
      key = option_key(trim(word))
      pos = pos + 1
      if (key == 0) cycle
      call get_command_argument(pos, o%value(key))
      pos = pos + 1
      if (echo_width(key) > 0) write(out, '(A,A)') echo_label(key)(1:echo_width(key)), trim(o%value(key))
      select case (key)
      case (K_OUTP)
        read(o%value(key), *) o%out_period
      case (K_OMP)
        read(o%value(key), *) o%threads
      case (K_PREC)
        o%precision = merge(PFM_MIXED, PFM_FP64, trim(o%value(key)) == 'mixed')
      end select
    enddo
  end subroutine read_cli

  !> The md() runs the options ask for: the rows of the list file (output
  !> prefix = out_path // the row's prefix), or the single settings file.
  subroutine plan_jobs(o, jobs)
    type(cli_options), intent(in) :: o
    type(md_job), allocatable, intent(out) :: jobs(:)
    integer :: u, n, k
    character(len=128) :: row_prefix

    if (len_trim(o%value(K_ILIST)) == 0) then
      allocate(jobs(1))
      jobs(1)%settings = o%value(K_IN)
      jobs(1)%prefix = o%value(K_PREFIX)
      return
    endif
    open(newunit=u, file=trim(o%value(K_IPATH))//trim(o%value(K_ILIST)), status='old', action='read')
    read(u, *) n
    allocate(jobs(max(n, 0)))
    do k = 1, size(jobs)
      read(u, *) jobs(k)%settings, row_prefix
      jobs(k)%prefix = trim(o%value(K_OPATH))//trim(row_prefix)
    enddo
    close(u)
  end subroutine plan_jobs

end module md_cli


program run_md_simulation
  use md_cli
  use pfm_md, only: md
  implicit none
  integer, parameter :: stdout = 6
  type(cli_options) :: opts
  type(md_job), allocatable :: jobs(:)
  logical :: listed
  integer :: k, sink

  write(stdout, '(A)') rule()
  call read_cli(opts, stdout)
  call plan_jobs(opts, jobs)
  ! a list of runs sends each run's summary to the all_out file
  listed = len_trim(opts%value(K_ILIST)) > 0
  sink = stdout
  if (listed) open(newunit=sink, file=trim(opts%value(K_OPATH))//trim(opts%value(K_OFILE)), action='write')
  do k = 1, size(jobs)
    write(stdout, '(A)') rule()
    if (listed) write(stdout, '(i6,A,A32,A32)') k, char(9), jobs(k)%settings, jobs(k)%prefix
    call md(stdout, sink, opts%value(K_IPATH), jobs(k)%settings, jobs(k)%prefix, opts%out_period, opts%threads, 1, &
            opts%precision)
    write(sink, *)
    write(stdout, '(A)') rule()
  enddo
  if (listed) close(sink)
  write(stdout, '(A)') rule()
end program run_md_simulation
