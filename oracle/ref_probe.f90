! ref_probe.f90 -- project-written probe driver for the reference MCM Fortran.
!
! TEST INFRASTRUCTURE ONLY.  This program is never part of the product path.
! It is linked against oracle/_ref/libmcmref.so, which build_ref.sh compiles in
! place from the read-only reference sources (MCM/*.f90); nothing from the
! reference is copied into this repository.
!
! What it does:
!   1. reads a binary deck (format written by mcmtop_amd/deck.py) and sets the
!      mcm_database module state exactly as the reference deck reader would
!      (mcm_control_v2.f90:29-235, mcm_matin.f90:35-208, mcm_nodein.f90:30-61,
!      mcm_inputvel.f90:28-131, mcm_inp_symmetry_v2.f90:111-137); the
!      material-constant set-up calls the reference's own mcm_sets1/3/4/10 and
!      mcm_printm (EOS V0 defaults), the deck text reader itself is bypassed
!      because it calls Intel time/ctime (mcm_print_version.f90:54-55);
!   2. runs the reference initialisation in the order of mcm_initial.f90:33-101
!      minus the output calls (:109-115);
!   3. runs the time-step body of mcm_solution.f90:41-116 minus the output and
!      termination checks, for a fixed number of steps;
!   4. writes raw little-endian dumps of the per-particle state after selected
!      stages (or, in timing mode, only the wall-clock of the step loop).
!
! Usage: ref_probe <deck.bin | deck.mcm> <outdir> <nsteps> <mode> [dump_steps...]
!   mode = 0 : timing only (no dumps, prints wall seconds for the step loop)
!   mode = 1 : dump every stage of the listed steps (and the init state)
!   mode = 2 : dump only the end-of-step state (after mcm_move) of listed steps
!   mode = 3 : like 1 for the listed steps, plus an energy line every step
!
program ref_probe
use mcm_database
use ref_deck
implicit none
character(len=512) :: deckfile, outdir, arg
integer :: nsteps, mode, nlist, istep, k, i, n
integer, allocatable :: dumplist(:)
integer(8) :: t0, t1, rate
logical :: dumpthis

call get_command_argument(1, deckfile)
call get_command_argument(2, outdir)
call get_command_argument(3, arg); read(arg,*) nsteps
call get_command_argument(4, arg); read(arg,*) mode
nlist = command_argument_count() - 4
allocate(dumplist(max(nlist,1)))
dumplist = -1
do k = 1, nlist
  call get_command_argument(4+k, arg); read(arg,*) dumplist(k)
enddo

open(unit=13, file=trim(outdir)//'/ref.log', status='replace', form='formatted')
mcm_openfiles = 0
mcm_openfiles(13) = 1

! a name ending in .mcm is the reference's text deck, read by its own card
! routines; anything else is the binary deck
n = len_trim(deckfile)
if (n .gt. 4 .and. deckfile(max(1,n-3):n) .eq. '.mcm') then
  call probe_read_text_deck(trim(deckfile))
else
  call probe_read_deck(trim(deckfile))
endif
call probe_initial()
if (mode.ne.0) call probe_dump(trim(outdir), 0, 0)

call system_clock(t0, rate)
do istep = 1, nsteps
  dumpthis = .false.
  if (mode.ne.0) then
    do k = 1, nlist
      if (dumplist(k).eq.istep) dumpthis = .true.
    enddo
  endif
  call probe_step(trim(outdir), istep, dumpthis .and. (mode.eq.1 .or. mode.eq.3))
  if (dumpthis .and. mode.eq.2) call probe_dump(trim(outdir), istep, 10)
  if (mode.eq.3) then
    write(*,'(a,i7,4es25.16)') 'E ', istep, mcm_ptime, mcm_dt, mcm_kinetice, mcm_internale
    flush(6)
  endif
enddo
call system_clock(t1)
write(*,'(a,i8,a,i8,a,es16.8)') 'probe: np=', mcm_np, ' steps=', nsteps, ' loop_seconds=', &
      real(t1-t0,8)/real(rate,8)
write(*,'(a,es25.16,a,es25.16)') 'probe: ptime=', mcm_ptime, ' dt=', mcm_dt
close(13)

contains


 ! initialisation sequence of mcm_initial.f90:33-101 (output calls omitted)
 subroutine probe_initial()
  integer :: i, l
  call mcm_init_boundcond
  call mcm_calcmass
  call mcm_initvbls
  call mcm_calcg
  call mcm_calcrhopc
  call mcm_initsigma
  call mcm_initrod
  call mcm_init_h
  mcm_timestep = 1
  do i = mcm_ssp, mcm_esp
    par(i)%tracerod = 0.0_d
    do l = 1, mcm_ndim
      par(i)%tracerod = par(i)%tracerod + par(i)%rod(l,l)
    enddo
    call mcm_aviscosity(i)
  enddo
  mcm_maxnbr = max(mcm_maxnbr, 160)
  mcm_g_maxnbr = max(mcm_g_maxnbr, 160)
  call mcm_init_neighb
  call mcm_calccritts
  if (mcm_itss .ne. 0.0_d) then
    mcm_dt = min(mcm_itss, mcm_critts*mcm_tssf)
  else
    mcm_dt = mcm_critts*mcm_tssf
  endif
  call mcm_calctotale
  call mcm_initold
 end subroutine probe_initial

 ! one pass of the loop body of mcm_solution.f90:41-126 (no output, no exit tests)
 subroutine probe_step(od, istep, dump)
  character(len=*) :: od
  integer :: istep
  logical :: dump
  ! mcm_ll_neighbours sets enough_mem once (mcm_ll_neighbours.f90:29) and never
  ! resets it for its second pass (:44,301-309), so any list-capacity growth
  ! after the first search is fatal, and the ghost capacity carries no margin
  ! (:83).  Capacities do not change list contents, so raise them before each
  ! search.
  mcm_maxnbr = max(mcm_maxnbr, 160)
  mcm_g_maxnbr = max(mcm_g_maxnbr, 160)
  ! The ghost array is sized once, count + 10% of N (mcm_init_boundary.f90:192),
  ! but the source layer fills as the bar moves onto the plane; the overflow
  ! retry keeps stale state (mcm_ghost_setup.f90:29,269-282) and aborts.  Ghosts
  ! are regenerated from scratch every search, so give the array room for
  ! 4N + 1024 entries between steps.
  if (mcm_boundary .and. mcm_max_ngp .lt. 4*mcm_np + 1024) then
    deallocate(gpar)
    mcm_max_ngp = 4*mcm_np + 1024
    allocate(gpar(mcm_max_ngp))
  endif
  call mcm_neighbours
  if (dump) call probe_dump(od, istep, 1)
  call mcm_strain
  if (dump) call probe_dump(od, istep, 2)
  if (mcm_h_opt.gt.0) call mcm_update_h
  if (dump) call probe_dump(od, istep, 3)
  call mcm_constitutive
  if (dump) call probe_dump(od, istep, 4)
  mcm_dtold = mcm_dt
  mcm_dt = mcm_tssf*mcm_critts
  mcm_dt = min(mcm_dt, 1.1_d*mcm_dtold)
  mcm_ptime = mcm_ptime + mcm_dt
  if (dump) call probe_dump(od, istep, 5)
  if (mcm_contacttype.gt.0) call mcm_rep_cont
  call mcm_momentum
  if (dump) call probe_dump(od, istep, 6)
  call mcm_boundcond
  if (dump) call probe_dump(od, istep, 7)
  call mcm_calctotale
  if (dump) call probe_dump(od, istep, 8)
  call mcm_update_velocity
  if (dump) call probe_dump(od, istep, 9)
  call mcm_move
  if (dump) call probe_dump(od, istep, 10)
  mcm_timestep = mcm_timestep + 1
 end subroutine probe_step


end program ref_probe
