! mcm_gpu_iface.f90 -- ISO_C_BINDING interface to libmcm_gpu.so (include/mcm_gpu.h).
!
! The module holds the context of the run that the same-named stage
! subroutines (mcm_gpu_stages.f90) act on, as the reference's step routines
! act on the module state of mcm_database.
module mcm_gpu_iface
  use iso_c_binding
  implicit none

  type, bind(c) :: mcm_gpu_options
    integer(c_int32_t) :: nbr_capacity = 0
    integer(c_int32_t) :: ghost_capacity = 0
    integer(c_int32_t) :: device = 0
    integer(c_int32_t) :: corrected_one_pass = 0
    integer(c_int32_t) :: pair_mapping = 0
    integer(c_int32_t) :: nbr_skin_milli = 0
    integer(c_int32_t) :: reserved(2) = 0
  end type

  type, bind(c) :: mcm_gpu_scalars
    real(c_double) :: dt, dtold, ptime, critts, hmax
    real(c_double) :: kinetice, internale, totale
    real(c_double) :: coord_min(3), coord_max(3)
    integer(c_int32_t) :: np, ngp, max_count, timestep
    integer(c_int32_t) :: error_flags, nummat
    integer(c_int32_t) :: skin_builds
    integer(c_int32_t) :: reserved(1)
  end type

  type(c_ptr), save :: mcm_gpu_context = c_null_ptr

  interface
    integer(c_int) function mcm_gpu_create_from_deck(path, opt, ctx) bind(c, name='mcm_gpu_create_from_deck')
      import :: c_int, c_char, c_ptr, mcm_gpu_options
      character(kind=c_char), dimension(*), intent(in) :: path
      type(mcm_gpu_options), intent(in) :: opt
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function mcm_gpu_convert_deck(path_in, path_out, msg, msg_len) bind(c, name='mcm_gpu_convert_deck')
      import :: c_int, c_char, c_int32_t
      character(kind=c_char), dimension(*), intent(in) :: path_in, path_out
      character(kind=c_char), dimension(*), intent(out) :: msg
      integer(c_int32_t), value :: msg_len
    end function
    subroutine mcm_gpu_destroy(ctx) bind(c, name='mcm_gpu_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function mcm_gpu_last_error(ctx) bind(c, name='mcm_gpu_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_initial(ctx) bind(c, name='mcm_gpu_initial')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_neighbours(ctx) bind(c, name='mcm_gpu_neighbours')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_strain(ctx) bind(c, name='mcm_gpu_strain')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_update_h(ctx) bind(c, name='mcm_gpu_update_h')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_constitutive(ctx) bind(c, name='mcm_gpu_constitutive')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_dt_update(ctx) bind(c, name='mcm_gpu_dt_update')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_set_time(ctx, dt, dtold, ptime) bind(c, name='mcm_gpu_set_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt, dtold, ptime
    end function
    integer(c_int) function mcm_gpu_get_energies(ctx, en, n) bind(c, name='mcm_gpu_get_energies')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: en
      integer(c_int32_t), value :: n
    end function
    ! per-particle state rows (MCM_STATE_WORDS = 86 per particle) and scalars
    integer(c_int) function mcm_gpu_upload_state(ctx, rows, np, s) bind(c, name='mcm_gpu_upload_state')
      import :: c_int, c_ptr, c_double, c_int64_t, mcm_gpu_scalars
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(in) :: rows
      integer(c_int64_t), value :: np
      type(mcm_gpu_scalars), intent(in) :: s
    end function
    ! the cap model's par%capa and par%ptype (upload 0: from the device)
    integer(c_int) function mcm_gpu_cap_state(ctx, capa, ptype, np, upload) bind(c, name='mcm_gpu_cap_state')
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(inout) :: capa
      integer(c_int32_t), dimension(*), intent(inout) :: ptype
      integer(c_int64_t), value :: np
      integer(c_int32_t), value :: upload
    end function
    integer(c_int) function mcm_gpu_download_state(ctx, rows, np, s) bind(c, name='mcm_gpu_download_state')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: rows
      integer(c_int64_t), value :: np
      type(c_ptr), value :: s
    end function
    ! mcm_gpu_download_neighbours with ids = NULL: the list lengths alone
    integer(c_int) function mcm_gpu_download_counts(ctx, counts, gcounts, ids, ids_len, cap) &
        bind(c, name='mcm_gpu_download_neighbours')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), dimension(*), intent(out) :: counts, gcounts
      type(c_ptr), value :: ids
      integer(c_int64_t), value :: ids_len
      integer(c_int32_t), intent(out) :: cap
    end function
    integer(c_int) function mcm_gpu_rep_cont(ctx) bind(c, name='mcm_gpu_rep_cont')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_momentum(ctx) bind(c, name='mcm_gpu_momentum')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_boundcond(ctx) bind(c, name='mcm_gpu_boundcond')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_calctotale(ctx) bind(c, name='mcm_gpu_calctotale')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_update_velocity(ctx) bind(c, name='mcm_gpu_update_velocity')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_move(ctx) bind(c, name='mcm_gpu_move')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mcm_gpu_calccritts(ctx, critts) bind(c, name='mcm_gpu_calccritts')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: critts
    end function
    integer(c_int) function mcm_gpu_step(ctx, nsteps) bind(c, name='mcm_gpu_step')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function mcm_gpu_get_scalars(ctx, s) bind(c, name='mcm_gpu_get_scalars')
      import :: c_int, c_ptr, mcm_gpu_scalars
      type(c_ptr), value :: ctx
      type(mcm_gpu_scalars), intent(out) :: s
    end function
    integer(c_int) function mcm_gpu_download(ctx, field, host, n) bind(c, name='mcm_gpu_download')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: field
      real(c_double), dimension(*), intent(out) :: host
      integer(c_int64_t), value :: n
    end function
    ! output sync before the host's writers: 17 packed values per particle
    ! (x, v, p, sigma 11 22 33 12 13 23, e, efps, rho, h), precision 8 or 4
    integer(c_int) function mcm_gpu_output_fields(ctx, precision, host, nbytes) bind(c, name='mcm_gpu_output_fields')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: precision
      type(c_ptr), value :: host
      integer(c_int64_t), value :: nbytes
    end function
    ! the state_opt 4 VTK file (material 0: all particles)
    integer(c_int) function mcm_gpu_write_vtk(ctx, path, material) bind(c, name='mcm_gpu_write_vtk')
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: material
    end function
    integer(c_int) function mcm_gpu_dump(ctx, path, step, stage) bind(c, name='mcm_gpu_dump')
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: step, stage
    end function
    function c_strlen(s) bind(c, name='strlen')
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: c_strlen
    end function
  end interface

contains

  ! nonzero status: print the library's message and stop, as the
  ! reference's fatal paths end in mcm_shutdown(2)
  subroutine mcm_gpu_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = mcm_gpu_last_error(mcm_gpu_context)
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    write(*,'(a,a,a,i0,a,*(a))') 'mcm_gpu: ', what, ' failed, status ', rc, ': ', msg
    stop 2
  end subroutine

  function c_string(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char) :: c(len_trim(s) + 1)
    integer :: i
    do i = 1, len_trim(s)
      c(i) = s(i:i)
    end do
    c(len_trim(s) + 1) = c_null_char
  end function
end module mcm_gpu_iface
