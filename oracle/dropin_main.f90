! dropin_main.f90 -- project-written driver of the reference's own
! initialisation and time loop (mcm_initial.f90, mcm_solution.f90), for the
! drop-in test of the GPU wrappers.
!
! TEST INFRASTRUCTURE ONLY.  oracle/build_ref.sh links it twice, both times
! with the reference's unchanged mcm_initial, mcm_init_neighb,
! mcm_neighbours, mcm_solution and output routines compiled in place:
!   oracle/_ref/ref_solution  with the reference's own step routines;
!   oracle/_ref/mcm_dropin    with mcmtop_amd/fortran/mcm_gpu_dropin.f90 in
!                             their place (the GPU library behind them).
! It stands in for main.f90's mcm_startup / mcm_getinput, which read the
! file names from the terminal and print the version banner through Intel's
! time/ctime (mcm_print_version.f90:54-55): the deck is read by ref_deck's
! probe_read_deck, and the output controls mcm_getinput would read come from
! the command line.
!
! Usage (in the directory that holds the deck; every output goes there):
!   dropin_main <deck> <endtime> <stpltime> <thpltime> <state_opt> [thnode ids...]
! <deck> is the binary deck's file name, at most 12 characters (mcm_filein).
! Outputs: the reference's writers' files with prefix 'run' (state_opt 1:
! EnSight; the time history run_thist.csv and one file per thnode), the log
! run.log, and final.bin: the state rows of par(:) when mcm_solution returns
! (after the last step's mcm_calctotale, mcm_solution.f90:77-101).
program dropin_main
use mcm_database
use ref_deck
implicit none
character(len=512) :: deckfile, arg
real(kind=real_acc) :: endtime, stpl, thpl
integer :: stopt, k, nth

call get_command_argument(1, deckfile)
call get_command_argument(2, arg); read(arg,*) endtime
call get_command_argument(3, arg); read(arg,*) stpl
call get_command_argument(4, arg); read(arg,*) thpl
call get_command_argument(5, arg); read(arg,*) stopt
nth = command_argument_count() - 5
if (len_trim(deckfile) .gt. 12) stop 'deck file name longer than 12 characters'

open(unit=13, file='run.log', status='replace', form='formatted')
mcm_openfiles = 0
mcm_openfiles(13) = 1
call probe_read_deck(trim(deckfile))

! what mcm_startup and mcm_getinput set besides the deck cards
mcm_filein = deckfile(1:12)
mcm_filelen(1) = len_trim(deckfile)
mcm_fileout = 'run'
mcm_filelen(2) = 3
mcm_filelen(3) = 3
mcm_title = 'drop-in test'
mcm_ctrlc = .false.
mcm_endtime = endtime
mcm_stpltime = stpl
mcm_thpltime = thpl
mcm_state_opt = stopt
mcm_nextsttime = 0.0_d
mcm_nextthtime = 0.0_d
mcm_istate = 0
mcm_thnode = nth
if (nth .gt. 0) then
  allocate(mcm_thnodeid(nth))
  do k = 1, nth
    call get_command_argument(5 + k, arg); read(arg,*) mcm_thnodeid(k)
  enddo
endif

call mcm_initial
! The reference sizes its ghost array once, to the initial count plus 10 % of
! N (mcm_init_boundary.f90:192), and its overflow retry aborts
! (mcm_ghost_setup.f90:269-282) once the bar has moved onto the plane; ghosts
! are made afresh by every search, so room for 4N + 1024 of them is given
! before the loop.  (The GPU wrappers keep no host ghosts.)
if (mcm_boundary) then
  if (allocated(gpar)) deallocate(gpar)
  mcm_max_ngp = 4*mcm_np + 1024
  allocate(gpar(mcm_max_ngp))
endif
call mcm_solution
call probe_dump_rows('final.bin', mcm_timestep, 8)
write(*,'(a,i8,a,i8,a,es25.16)') 'dropin: np=', mcm_np, ' timestep=', mcm_timestep, ' ptime=', mcm_ptime
close(13)
end program dropin_main
