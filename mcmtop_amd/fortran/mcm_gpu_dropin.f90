! mcm_gpu_dropin.f90 -- the reference's time-step subroutines under their own
! names, for a host built from the reference's own module and drivers
! (mcm_database.f90, mcm_initial.f90, mcm_init_neighb.f90, mcm_neighbours.f90,
! mcm_solution.f90).  Linked in place of the reference's objects of the same
! names, they run each stage on the GPU and keep the host's module state
! (mcm_database) in step with the device where the host reads it:
!
!   * the first search (mcm_ll_neighbours from mcm_init_neighb) creates the
!     device run from the deck the host read (mcm_filein) and uploads the
!     state the host's initialisation (mcm_initial.f90:33-78) made in par(:);
!   * the host's mcm_dt, mcm_dtold and mcm_ptime, which mcm_solution.f90:58-62
!     computes, go to the device before the search and before the momentum
!     stage; mcm_critts comes back after mcm_constitutive and mcm_calccritts;
!   * the energies of mcm_calctotale come back every step, and par(:) with
!     the neighbour counts is downloaded when mcm_solution is about to write
!     output (its tests at mcm_solution.f90:82,89 and the initial output of
!     mcm_initial.f90:109-115);
!   * WriteVTK / WriteVTKmat (mcm_VTK.f90, state_opt 4) write the same files
!     from the device state.
!
! The search's binning, boundary set-up and ghost generation happen inside
! the device search, so mcm_setuplist, mcm_init_boundary and mcm_ghost_setup
! do nothing here, and mcm_rhoupd is part of the fused constitutive update.
! Neighbour lists and ghosts stay on the device.
module mcm_gpu_dropin_sync
  use iso_c_binding
  use mcm_database
  use mcm_gpu_iface
  implicit none
  integer, parameter :: nwords = 86   ! MCM_STATE_WORDS

contains

  ! par(:) -> state rows (the row layout of include/mcm_gpu.h)
  subroutine pack_par(rows)
    real(c_double), intent(out) :: rows(nwords, mcm_np)
    integer :: i
    do i = 1, mcm_np
      rows(1:3, i) = par(i)%x
      rows(4:6, i) = par(i)%v
      rows(7:9, i) = par(i)%a
      rows(10, i) = par(i)%rho;      rows(11, i) = par(i)%rhoold
      rows(12, i) = par(i)%h;        rows(13, i) = par(i)%hold
      rows(14, i) = par(i)%p;        rows(15, i) = par(i)%e
      rows(16, i) = par(i)%etry;     rows(17, i) = par(i)%einc
      rows(18, i) = par(i)%c;        rows(19, i) = par(i)%vabs
      rows(20, i) = par(i)%tracerod
      rows(21, i) = par(i)%pcut;     rows(22, i) = par(i)%p_cut
      rows(23, i) = par(i)%efps;     rows(24, i) = par(i)%fail
      rows(25, i) = par(i)%mindist;  rows(26, i) = par(i)%critts
      rows(27, i) = par(i)%mass;     rows(28, i) = par(i)%rho0
      rows(29:37, i) = reshape(par(i)%sigma, (/9/))
      rows(38:46, i) = reshape(par(i)%s, (/9/))
      rows(47:55, i) = reshape(par(i)%alfa, (/9/))
      rows(56:64, i) = reshape(par(i)%q, (/9/))
      rows(65:73, i) = reshape(par(i)%rod, (/9/))
      rows(74:82, i) = reshape(par(i)%spin, (/9/))
      rows(83, i) = par(i)%h0
      rows(84:86, i) = par(i)%xzero
    end do
  end subroutine

  ! state rows -> par(:)
  subroutine unpack_par(rows)
    real(c_double), intent(in) :: rows(nwords, mcm_np)
    integer :: i
    do i = 1, mcm_np
      par(i)%x = rows(1:3, i)
      par(i)%v = rows(4:6, i)
      par(i)%a = rows(7:9, i)
      par(i)%rho = rows(10, i);      par(i)%rhoold = rows(11, i)
      par(i)%h = rows(12, i);        par(i)%hold = rows(13, i)
      par(i)%p = rows(14, i);        par(i)%e = rows(15, i)
      par(i)%etry = rows(16, i);     par(i)%einc = rows(17, i)
      par(i)%c = rows(18, i);        par(i)%vabs = rows(19, i)
      par(i)%tracerod = rows(20, i)
      par(i)%pcut = rows(21, i);     par(i)%p_cut = rows(22, i)
      par(i)%efps = rows(23, i);     par(i)%fail = rows(24, i)
      par(i)%mindist = rows(25, i);  par(i)%critts = rows(26, i)
      par(i)%mass = rows(27, i);     par(i)%rho0 = rows(28, i)
      par(i)%sigma = reshape(rows(29:37, i), (/3, 3/))
      par(i)%s = reshape(rows(38:46, i), (/3, 3/))
      par(i)%alfa = reshape(rows(47:55, i), (/3, 3/))
      par(i)%q = reshape(rows(56:64, i), (/3, 3/))
      par(i)%rod = reshape(rows(65:73, i), (/3, 3/))
      par(i)%spin = reshape(rows(74:82, i), (/3, 3/))
      par(i)%h0 = rows(83, i)
      par(i)%xzero = rows(84:86, i)
    end do
  end subroutine

  subroutine host_scalars(s)
    type(mcm_gpu_scalars), intent(out) :: s
    integer :: k
    s%dt = mcm_dt
    s%dtold = mcm_dtold
    s%ptime = mcm_ptime
    s%critts = mcm_critts
    s%hmax = mcm_hmax
    do k = 1, 3
      s%coord_min(k) = mcm_coord_maxmin(1, k)
      s%coord_max(k) = mcm_coord_maxmin(2, k)
    end do
    s%timestep = mcm_timestep
    s%kinetice = 0; s%internale = 0; s%totale = 0
    s%np = mcm_np; s%ngp = 0; s%max_count = 0; s%error_flags = 0; s%nummat = mcm_nummat
    s%skin_builds = 0; s%reserved = 0
  end subroutine

  ! the device run, made on first use from the deck the host read, with the
  ! state of the host's initialisation
  subroutine dropin_attach()
    type(mcm_gpu_options) :: opt
    type(mcm_gpu_scalars) :: s
    real(c_double), allocatable :: rows(:,:)
    if (c_associated(mcm_gpu_context)) return
    call mcm_gpu_check(mcm_gpu_create_from_deck(c_string(mcm_filein(1:mcm_filelen(1))), opt, mcm_gpu_context), &
                       'create (deck '//mcm_filein(1:mcm_filelen(1))//')')
    allocate(rows(nwords, mcm_np))
    call pack_par(rows)
    call host_scalars(s)
    call mcm_gpu_check(mcm_gpu_upload_state(mcm_gpu_context, rows, int(mcm_np, c_int64_t), s), 'upload_state')
    call cap_state(1_c_int32_t)
  end subroutine

  ! par%capa and par%ptype (the cap model's, mcm_database.f90:195-196)
  ! between host and device: upload 1 to the device, 0 from it
  subroutine cap_state(upload)
    integer(c_int32_t), intent(in) :: upload
    real(c_double), allocatable :: capa(:)
    integer(c_int32_t), allocatable :: ptype(:)
    integer :: i
    allocate(capa(mcm_np), ptype(mcm_np))
    if (upload == 1) then
      do i = 1, mcm_np
        capa(i) = par(i)%capa
        ptype(i) = par(i)%ptype
      end do
    end if
    call mcm_gpu_check(mcm_gpu_cap_state(mcm_gpu_context, capa, ptype, int(mcm_np, c_int64_t), upload), 'cap_state')
    if (upload == 0) then
      do i = 1, mcm_np
        par(i)%capa = capa(i)
        par(i)%ptype = ptype(i)
      end do
    end if
  end subroutine

  ! the host's time step (mcm_solution.f90:58-62) to the device
  subroutine push_time()
    call mcm_gpu_check(mcm_gpu_set_time(mcm_gpu_context, mcm_dt, mcm_dtold, mcm_ptime), 'set_time')
  end subroutine

  subroutine pull_scalars()
    type(mcm_gpu_scalars) :: s
    integer :: k
    call mcm_gpu_check(mcm_gpu_get_scalars(mcm_gpu_context, s), 'get_scalars')
    mcm_critts = s%critts
    mcm_hmax = s%hmax
    mcm_ngp = s%ngp
    do k = 1, 3
      mcm_coord_maxmin(1, k) = s%coord_min(k)
      mcm_coord_maxmin(2, k) = s%coord_max(k)
    end do
  end subroutine

  ! energies of mcm_calctotale into the module (mcm_calctotale.f90:33-59)
  subroutine pull_energies()
    real(c_double) :: en(5 * mcm_nummat)
    integer :: m
    call mcm_gpu_check(mcm_gpu_get_energies(mcm_gpu_context, en, int(5 * mcm_nummat, c_int32_t)), 'get_energies')
    mcm_kinetice = 0.0_d
    mcm_internale = 0.0_d
    do m = 1, mcm_nummat
      mcm_mat(m)%ke = en(5 * m - 4)
      mcm_mat(m)%ie = en(5 * m - 3)
      mcm_mat(m)%mom = en(5 * m - 2:5 * m)
      mcm_kinetice = mcm_kinetice + mcm_mat(m)%ke
      mcm_internale = mcm_internale + mcm_mat(m)%ie
    end do
    mcm_totale = mcm_kinetice + mcm_internale
  end subroutine

  ! par(:) and the neighbour counts from the device, for the host's writers
  subroutine pull_state()
    real(c_double), allocatable :: rows(:,:)
    integer(c_int32_t), allocatable :: nn(:), gn(:)
    integer(c_int32_t) :: cap
    integer :: i
    allocate(rows(nwords, mcm_np), nn(mcm_np), gn(mcm_np))
    call mcm_gpu_check(mcm_gpu_download_state(mcm_gpu_context, rows, int(mcm_np, c_int64_t), c_null_ptr), &
                       'download_state')
    call unpack_par(rows)
    call cap_state(0_c_int32_t)
    call mcm_gpu_check(mcm_gpu_download_counts(mcm_gpu_context, nn, gn, c_null_ptr, 0_c_int64_t, cap), &
                       'download_neighbours')
    do i = 1, mcm_np
      par(i)%nnbr = nn(i)
      par(i)%g_nnbr = gn(i)
    end do
  end subroutine
end module mcm_gpu_dropin_sync

! mcm_ll_neighbours.f90:1-339 with setuplist and ghost_setup: the device
! search.  ip_flag 1 (self in the list) is the corrected discretisation.
subroutine mcm_ll_neighbours(ip_flag)
  use mcm_gpu_dropin_sync
  integer :: ip_flag
  call dropin_attach()
  if (ip_flag .ne. mcm_disctype) then
    write(*,'(a,i0,a,i0)') 'mcm_gpu: mcm_ll_neighbours(', ip_flag, ') on a run with disctype ', mcm_disctype
    stop 2
  end if
  call push_time()
  call mcm_gpu_check(mcm_gpu_neighbours(mcm_gpu_context), 'mcm_ll_neighbours')
  call pull_scalars()
end subroutine

! binning (mcm_setuplist.f90), plane set-up (mcm_init_boundary.f90) and
! ghost generation (mcm_ghost_setup.f90) are done by the device search
subroutine mcm_setuplist
end subroutine

subroutine mcm_init_boundary
end subroutine

subroutine mcm_ghost_setup
end subroutine

subroutine mcm_strain
  use mcm_gpu_dropin_sync
  call mcm_gpu_check(mcm_gpu_strain(mcm_gpu_context), 'mcm_strain')
end subroutine

subroutine mcm_update_h
  use mcm_gpu_dropin_sync
  call mcm_gpu_check(mcm_gpu_update_h(mcm_gpu_context), 'mcm_update_h')
end subroutine

! mcm_rhoupd.f90 is the first step of the fused constitutive update
subroutine mcm_rhoupd
end subroutine

! mcm_constitutive.f90:26-87 (rhoupd, models, calccritts); mcm_critts back
! to the host for mcm_solution.f90:59
subroutine mcm_constitutive
  use mcm_gpu_dropin_sync
  call mcm_gpu_check(mcm_gpu_constitutive(mcm_gpu_context), 'mcm_constitutive')
  call pull_scalars()
end subroutine

subroutine mcm_calccritts
  use mcm_gpu_dropin_sync
  call mcm_gpu_check(mcm_gpu_calccritts(mcm_gpu_context, mcm_critts), 'mcm_calccritts')
end subroutine

subroutine mcm_rep_cont
  use mcm_gpu_dropin_sync
  call push_time()
  call mcm_gpu_check(mcm_gpu_rep_cont(mcm_gpu_context), 'mcm_rep_cont')
end subroutine

! mcm_momentum.f90:22-43 with the host's new dt.  The per-material contact
! force totals that mcm_oldacceleration.f90:36-56 keeps for mcm_timehist are
! zero without penalty contact; with it they are not carried over.
subroutine mcm_momentum
  use mcm_gpu_dropin_sync
  integer :: m
  call push_time()
  call mcm_gpu_check(mcm_gpu_momentum(mcm_gpu_context), 'mcm_momentum')
  if (mcm_disctype .eq. 0) then
    do m = 1, mcm_nummat
      mcm_mat(m)%xcf = 0.0
      mcm_mat(m)%ycf = 0.0
      mcm_mat(m)%zcf = 0.0
    end do
  end if
end subroutine

subroutine mcm_boundcond
  use mcm_gpu_dropin_sync
  call mcm_gpu_check(mcm_gpu_boundcond(mcm_gpu_context), 'mcm_boundcond')
end subroutine

! mcm_calctotale.f90:27-60; par(:) follows when the caller writes output next
! (the state and time-history tests of mcm_solution.f90:82,89, which hold at
! the initial output too)
subroutine mcm_calctotale
  use mcm_gpu_dropin_sync
  call mcm_gpu_check(mcm_gpu_calctotale(mcm_gpu_context), 'mcm_calctotale')
  call pull_energies()
  if (mcm_ptime .ge. mcm_nextsttime .or. mcm_ptime .ge. mcm_endtime .or. mcm_ptime .ge. mcm_nextthtime &
      .or. mcm_timestep .le. 1) call pull_state()
end subroutine

subroutine mcm_update_velocity
  use mcm_gpu_dropin_sync
  call mcm_gpu_check(mcm_gpu_update_velocity(mcm_gpu_context), 'mcm_update_velocity')
end subroutine

subroutine mcm_move
  use mcm_gpu_dropin_sync
  call mcm_gpu_check(mcm_gpu_move(mcm_gpu_context), 'mcm_move')
  call pull_scalars()
end subroutine

! mcm_VTK.f90 (state_opt 4): <fileout><istate>.vtk and, per material,
! <mat><fileout><istate>.vtk, written from the device state
subroutine WriteVTK
  use mcm_gpu_dropin_sync
  character(len=25) :: name
  write(name, '(a,i0,a)') mcm_fileout(1:mcm_filelen(2)), mcm_istate, '.vtk'
  call mcm_gpu_check(mcm_gpu_write_vtk(mcm_gpu_context, c_string(name), 0_c_int32_t), 'WriteVTK')
end subroutine

subroutine WriteVTKmat(matID)
  use mcm_gpu_dropin_sync
  integer :: matID
  character(len=25) :: name
  write(name, '(i1,a,i0,a)') matID, mcm_fileout(1:mcm_filelen(2)), mcm_istate, '.vtk'
  call mcm_gpu_check(mcm_gpu_write_vtk(mcm_gpu_context, c_string(name), int(matID, c_int32_t)), 'WriteVTKmat')
end subroutine
