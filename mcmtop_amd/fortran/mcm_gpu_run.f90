! mcm_gpu_run.f90 -- Fortran host driver of the GPU step.
!
! Usage: mcm_gpu_run <deck> <nsteps> [fused|stages] [dump.bin]
! Creates the device run from a deck (the reference's text deck or the
! binary deck), runs the initialisation of
! mcm_initial.f90:33-101 on the device, then either the loop body of
! mcm_solution.f90:41-126 through the same-named stage subroutines
! ("stages", dt and ptime advanced on the device between mcm_constitutive
! and mcm_momentum as at :58-62) or the fused captured step ("fused").
program mcm_gpu_run
  use iso_c_binding
  use mcm_gpu_iface
  implicit none
  character(len=512) :: deck, arg, mode, dumpfile
  integer :: nsteps, istep
  type(mcm_gpu_options) :: opt
  type(mcm_gpu_scalars) :: s

  call get_command_argument(1, deck)
  call get_command_argument(2, arg)
  read(arg, *) nsteps
  mode = 'fused'
  if (command_argument_count() >= 3) call get_command_argument(3, mode)
  dumpfile = ''
  if (command_argument_count() >= 4) call get_command_argument(4, dumpfile)

  call mcm_gpu_check(mcm_gpu_create_from_deck(c_string(deck), opt, mcm_gpu_context), 'create')
  call mcm_gpu_check(mcm_gpu_initial(mcm_gpu_context), 'initial')
  if (trim(mode) == 'stages') then
    do istep = 1, nsteps
      call mcm_neighbours
      call mcm_strain
      call mcm_update_h
      call mcm_constitutive
      call mcm_gpu_check(mcm_gpu_dt_update(mcm_gpu_context), 'dt update')
      call mcm_rep_cont        ! mcm_solution.f90:64 (does nothing without penalty contact)
      call mcm_momentum
      call mcm_boundcond
      call mcm_calctotale
      call mcm_update_velocity
      call mcm_move
    end do
  else
    call mcm_gpu_check(mcm_gpu_step(mcm_gpu_context, int(nsteps, c_int32_t)), 'step')
  end if
  call mcm_gpu_check(mcm_gpu_get_scalars(mcm_gpu_context, s), 'scalars')
  write(*,'(a,i0,a,i0,3(a,es24.16))') 'mcm_gpu_run: np=', s%np, ' steps=', nsteps, ' ptime=', s%ptime, &
        ' ke=', s%kinetice, ' ie=', s%internale
  if (len_trim(dumpfile) > 0) then
    call mcm_gpu_check(mcm_gpu_dump(mcm_gpu_context, c_string(dumpfile), int(nsteps, c_int32_t), 10_c_int32_t), &
                       'dump')
  end if
  call mcm_gpu_destroy(mcm_gpu_context)
end program mcm_gpu_run
