! mcm_gpu_stages.f90 -- argument-less step subroutines with the reference's
! names (MCM/mcm_solution.f90:41-116 calls them), each one a call into the
! GPU library on the module context.  Linked in place of the reference's
! hot-path objects they make the caller's loop a drop-in.
subroutine mcm_neighbours
  use mcm_gpu_iface
  call mcm_gpu_check(mcm_gpu_neighbours(mcm_gpu_context), 'mcm_neighbours')
end subroutine

subroutine mcm_strain
  use mcm_gpu_iface
  call mcm_gpu_check(mcm_gpu_strain(mcm_gpu_context), 'mcm_strain')
end subroutine

subroutine mcm_update_h
  use mcm_gpu_iface
  call mcm_gpu_check(mcm_gpu_update_h(mcm_gpu_context), 'mcm_update_h')
end subroutine

! mcm_constitutive includes rhoupd and calccritts on the device
subroutine mcm_constitutive
  use mcm_gpu_iface
  call mcm_gpu_check(mcm_gpu_constitutive(mcm_gpu_context), 'mcm_constitutive')
end subroutine

subroutine mcm_rep_cont
  use mcm_gpu_iface
  call mcm_gpu_check(mcm_gpu_rep_cont(mcm_gpu_context), 'mcm_rep_cont')
end subroutine

subroutine mcm_momentum
  use mcm_gpu_iface
  call mcm_gpu_check(mcm_gpu_momentum(mcm_gpu_context), 'mcm_momentum')
end subroutine

subroutine mcm_boundcond
  use mcm_gpu_iface
  call mcm_gpu_check(mcm_gpu_boundcond(mcm_gpu_context), 'mcm_boundcond')
end subroutine

subroutine mcm_calctotale
  use mcm_gpu_iface
  call mcm_gpu_check(mcm_gpu_calctotale(mcm_gpu_context), 'mcm_calctotale')
end subroutine

subroutine mcm_update_velocity
  use mcm_gpu_iface
  call mcm_gpu_check(mcm_gpu_update_velocity(mcm_gpu_context), 'mcm_update_velocity')
end subroutine

subroutine mcm_move
  use mcm_gpu_iface
  call mcm_gpu_check(mcm_gpu_move(mcm_gpu_context), 'mcm_move')
end subroutine
