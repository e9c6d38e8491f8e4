! ref_deck.f90 -- deck reading and state dumps shared by the project-written
! reference drivers (ref_probe.f90, dropin_main.f90).
!
! TEST INFRASTRUCTURE ONLY.  Compiled against the module files of the
! reference build (oracle/build_ref.sh); nothing from the reference is copied.
!
! probe_read_deck reads a binary deck (mcmtop_amd/deck.py) and sets the
! mcm_database module state as the reference deck reader would
! (mcm_control_v2.f90:29-235, mcm_matin.f90:35-208, mcm_nodein.f90:30-61,
! mcm_inputvel.f90:28-131, mcm_inp_symmetry_v2.f90:111-137), calling the
! reference's own material set-up routines.  probe_read_text_deck reads the
! reference's version-2 text deck with the reference's own card routines
! (mcm_getinput.f90:89-127: mcm_matin, mcm_nodein, mcm_inputvel,
! mcm_inp_symmetry_v2, mcm_contact_in, mcm_baseaccel_in), restating only the
! control cards.  probe_dump writes the state dump read by oracle/dumpio.py.
module ref_deck
use mcm_database
implicit none
contains

 subroutine probe_read_deck(fname)
  character(len=*) :: fname
  integer :: magic, u, m, n, j, ios, tag
  integer :: mirr_opt, sym_flag(6), nst
  integer :: model, eos, visc
  real(kind=real_acc) :: rho, avq, avl, mass, h, rmin, rmax
  real(kind=real_acc) :: card(48), eoscard(24), prop(48), symx(6)
  real(kind=real_acc), allocatable :: xin(:,:), vin(:,:), pm(:), ph(:), pr(:), pe(:)
  integer, allocatable :: matin(:), bcin(:)

  open(newunit=u, file=fname, access='stream', form='unformatted', status='old')
  read(u) magic
  if (magic .ne. 1296256305) stop 'bad deck magic'
  read(u) mcm_np, mcm_nummat, mcm_disctype, mcm_h_opt, mcm_tcrit_opt, mcm_contacttype, &
          mcm_krtype, mcm_maxnbr, mcm_massopt, mcm_init_h_opt, mcm_init_rhoe, mirr_opt, sym_flag, nst
  read(u) mcm_tssf, mcm_itss, mcm_endtime, symx

  ! control card defaults and derived values (mcm_control_v2.f90:29-235)
  pi = acos(-1.0_d)
  mcm_axopt = 3
  mcm_ndim = 3
  mcm_max_np = mcm_np
  mcm_state_opt = 0
  mcm_thnode = 0
  mcm_num_transducer = 0
  mcm_status_interval = 100
  mcm_veloc_opt = nst
  mcm_init_v_opt = 1
  mcm_nlcur = 0
  mcm_nthpx = 0; mcm_nthpy = 0; mcm_nthpz = 0
  mcm_drelax = .false.
  mcm_drelax_scale = 0.0_d
  mcm_baseaccel = .false.
  mcm_boundary = (mirr_opt .eq. 1)
  if (mcm_tssf.eq.0.0_d) mcm_tssf = 0.8_d
  if (mcm_krtype.eq.0) mcm_krtype = 1
  if (mcm_maxnbr.eq.0) mcm_maxnbr = 40
  mcm_maxcont = mcm_maxnbr
  mcm_g_maxnbr = mcm_maxnbr/2
  mcm_g_maxcont = mcm_g_maxnbr
  mcm_ssp = 1; mcm_svp = 1; mcm_esp = mcm_np; mcm_evp = mcm_np
  mcm_ptime = 0.0_d
  mcm_timestep = 0

  call mcm_allocate_memory
  call mcm_initialise_database
  ! mcm_initialise_database leaves the cap model's capa and ptype unset, and
  ! the model's first update reads capa (mcm_f3dm25.f90:162); a freshly
  ! mapped allocation holds zeros, which is set here explicitly
  par(:)%capa = 0.0_d
  par(:)%ptype = 0

  do m = 1, mcm_nummat
    mcm_mat(m)%model = 0
    mcm_mat(m)%eos = 0
  enddo
  do m = 1, mcm_nummat
    read(u) model, eos, visc
    read(u) rho, avq, avl, mass, h, rmin, rmax
    read(u) card
    read(u) eoscard
    n = m
    mcm_mat(n)%model = model
    mcm_mat(n)%eos = eos
    mcm_mat(n)%visc_type = visc
    mcm_mat(n)%rho = rho
    mcm_mat(n)%av_q = avq
    mcm_mat(n)%av_l = avl
    ! material card defaults (mcm_matin.f90:46-56)
    if (mcm_mat(n)%visc_type.eq.0) mcm_mat(n)%visc_type = 1
    if (mcm_mat(n)%av_l.eq.0.0) mcm_mat(n)%av_l = 0.06_d
    if (mcm_mat(n)%av_q.eq.0.0) mcm_mat(n)%av_q = 1.5_d
    mcm_mat(n)%mass = mass
    mcm_mat(n)%h = h
    mcm_mat(n)%rho_min = rmin
    mcm_mat(n)%rho_max = rmax
    if (mcm_mat(n)%rho_min.eq.0.0_d) mcm_mat(n)%rho_min = 1.0E-10_d
    if (mcm_mat(n)%rho_max.eq.0.0_d) mcm_mat(n)%rho_max = 1.0E+10_d
    mcm_mat(n)%strinput = card
    mcm_mat(n)%strinput2 = 0.0_d
    mcm_mat(n)%eosinput = eoscard
    prop = 0.0_d
    select case (model)
    case(1)
      call mcm_sets1(mcm_mat(n)%strinput(1), prop)
    case(3)
      call mcm_sets3(mcm_mat(n)%strinput(1), prop)
    case(4)
      call mcm_sets4(mcm_mat(n)%strinput(1), prop)
    case(10)
      call mcm_sets10(mcm_mat(n)%strinput(1), prop)
    end select
    call mcm_printm(n, prop)
  enddo

  allocate(xin(3,mcm_np), vin(3,mcm_np), matin(mcm_np), bcin(mcm_np))
  allocate(pm(mcm_np), ph(mcm_np), pr(mcm_np), pe(mcm_np))
  read(u) xin
  read(u) vin
  read(u) matin
  read(u) bcin
  read(u) pm, ph, pr, pe
  ! optional penalty-contact section (mcm_contact_in.f90 cards): 'CONT',
  ! ncontmats, cont_opt, then K and n per material
  if (mcm_contacttype.gt.0) then
    read(u, iostat=ios) tag
    if (ios.ne.0 .or. tag.ne.1129270868) stop 'contact deck without contact section'
    read(u) mcm_ncontmats, mcm_cont_opt
    do m = 1, mcm_nummat
      read(u) mcm_k_cont(m), mcm_n_cont(m)
    enddo
  endif
  ! optional 'OPTS' section: nthp(3), drelax_scale, base_a(3)
  ! (mcm_control_v2.f90:96,112,223-243 and mcm_baseaccel_in.f90:25-36)
  read(u, iostat=ios) tag
  if (ios.eq.0 .and. tag.eq.1330664531) then
    read(u) mcm_nthpx, mcm_nthpy, mcm_nthpz
    read(u) mcm_drelax_scale, mcm_base_a
    if (mcm_drelax_scale.gt.0.1_d.and.mcm_drelax_scale.lt.1.0_d) mcm_drelax = .true.
    if (mcm_nthpx.eq.1.or.mcm_nthpy.eq.1.or.mcm_nthpz.eq.1) mcm_baseaccel = .true.
    if (mcm_nthpx.ne.1) mcm_base_a(1) = 0.0_d
    if (mcm_nthpy.ne.1) mcm_base_a(2) = 0.0_d
    if (mcm_nthpz.ne.1) mcm_base_a(3) = 0.0_d
  endif
  close(u)
  do j = 1, mcm_np
    par(j)%x = xin(:,j)
    par(j)%v = vin(:,j)
    par(j)%mat = matin(j)
    par(j)%dispbc = bcin(j)
    if (mcm_massopt.eq.1) par(j)%mass = pm(j)
    if (mcm_init_h_opt.eq.1) par(j)%h = ph(j)
    if (mcm_init_rhoe.eq.1) then
      par(j)%rho = pr(j)
      par(j)%e = pe(j)
    endif
  enddo

  ! boundary planes (mcm_inp_symmetry_v2.f90:111-137)
  mcm_boundary_code(1,1) = sym_flag(1); mcm_boundary_code(2,1) = sym_flag(2)
  mcm_boundary_code(1,2) = sym_flag(3); mcm_boundary_code(2,2) = sym_flag(4)
  mcm_boundary_code(1,3) = sym_flag(5); mcm_boundary_code(2,3) = sym_flag(6)
  mcm_boundary_x(1,1) = symx(1); mcm_boundary_x(2,1) = symx(2)
  mcm_boundary_x(1,2) = symx(3); mcm_boundary_x(2,2) = symx(4)
  mcm_boundary_x(1,3) = symx(5); mcm_boundary_x(2,3) = symx(6)
  if (mcm_boundary) then
    if (all(sym_flag.eq.0)) mcm_boundary = .false.
  else
    mcm_boundary_code = 0
  endif
  do j = 1, 3
    mcm_boundary_type(j) = 1
    if (mcm_boundary_code(1,j).gt.0) mcm_boundary_type(j) = mcm_boundary_type(j) + 1
    if (mcm_boundary_code(2,j).gt.0) mcm_boundary_type(j) = mcm_boundary_type(j) + 2
  enddo
 end subroutine probe_read_deck

 ! The version-2 text deck through the reference's own card routines, in the
 ! order of mcm_getinput.f90:89-127.  Only mcm_control_v2 is restated (its
 ! defaults, the six control cards with the reference's formats, the seventh
 ! card skipped, the derived values; mcm_control_v2.f90:29-235): the routine
 ! itself also opens the run's log and prints the version banner through
 ! Intel's time/ctime (mcm_print_version.f90:54-55), which flang lacks.
 subroutine probe_read_text_deck(fname)
  character(len=*) :: fname
  character(len=80) :: txts
  integer :: lcount, vern, mirr_opt, keep_files
  open(unit=101, file=fname, status='old')
  mcm_openfiles(1) = 1
  mcm_boundary_type = 1
  mcm_boundary_code = 0
  call mcm_gttxsg(txts, lcount)
  read(unit=txts, fmt='(a78,i2)') mcm_title, vern
  if (vern .ne. 2) stop 'text deck: input format version 2 expected'
  pi = acos(-1.0_d)
  mcm_disctype = 0; mcm_max_np = 0; mcm_itss = 0.0_d; mcm_state_opt = 0; mcm_thnode = 0
  mcm_num_transducer = 0; mcm_status_interval = 100; mcm_restart_interval = 0; mcm_run_restart = 0
  keep_files = 0; mcm_veloc_opt = 0; mcm_massopt = 0; mcm_init_h_opt = 0; mcm_init_rhoe = 0
  mcm_contacttype = 0; mcm_tcrit_opt = 0; mcm_boundary = .false.; mirr_opt = 0; mcm_nlcur = 0
  mcm_h_opt = 0; mcm_krtype = 1; mcm_maxnbr = 40; mcm_nthpx = 0; mcm_nthpy = 0; mcm_nthpz = 0
  mcm_drelax = .false.
  call mcm_gttxsg(txts, lcount)
  read(unit=txts, fmt='(2i5,5i10)') mcm_axopt, mcm_disctype, mcm_nummat, mcm_np, mcm_max_np, mcm_nstressp, &
       mcm_nvelocp
  call mcm_gttxsg(txts, lcount)
  read(unit=txts, fmt='(4e10.0)') mcm_endtime, mcm_tssf, mcm_itss, mcm_drelax_scale
  call mcm_gttxsg(txts, lcount)
  read(unit=txts, fmt='(e10.0,i5,e10.0,i10,i5,i10,2i5)') mcm_stpltime, mcm_state_opt, mcm_thpltime, mcm_thnode, &
       mcm_num_transducer, mcm_status_interval, mcm_restart_interval, mcm_run_restart
  call mcm_gttxsg(txts, lcount)
  read(unit=txts, fmt='(5i5)') keep_files, mcm_init_v_opt, mcm_massopt, mcm_init_h_opt, mcm_init_rhoe
  call mcm_gttxsg(txts, lcount)
  read(unit=txts, fmt='(10i5)') mcm_contacttype, mcm_tcrit_opt, mcm_veloc_opt, mirr_opt, mcm_nlcur, mcm_nthpx, &
       mcm_nthpy, mcm_nthpz
  call mcm_gttxsg(txts, lcount)
  read(unit=txts, fmt='(3i5)') mcm_h_opt, mcm_krtype, mcm_maxnbr
  call mcm_gttxsg(txts, lcount)
  if (mcm_tssf .eq. 0.0_d) mcm_tssf = 0.8_d
  if (mcm_krtype .eq. 0) mcm_krtype = 1
  if (mcm_maxnbr .eq. 0) mcm_maxnbr = 40
  if (mcm_status_interval .lt. 1) mcm_status_interval = 100
  if (mirr_opt .eq. 1) mcm_boundary = .true.
  mcm_maxcont = mcm_maxnbr
  mcm_g_maxnbr = mcm_maxnbr/2
  mcm_g_maxcont = mcm_g_maxnbr
  if (mcm_max_np .lt. mcm_np) mcm_max_np = mcm_np
  if (mcm_axopt .lt. 1 .or. mcm_axopt .gt. 3) stop 'text deck: axis option 1, 2 or 3 expected'
  mcm_ndim = mcm_axopt
  if (mcm_drelax_scale .gt. 0.1_d .and. mcm_drelax_scale .lt. 1.0_d) mcm_drelax = .true.
  mcm_ssp = 1; mcm_svp = 1; mcm_esp = mcm_np; mcm_evp = mcm_np
  if (mcm_nthpx .eq. 1 .or. mcm_nthpy .eq. 1 .or. mcm_nthpz .eq. 1) mcm_baseaccel = .true.
  mcm_ptime = 0.0_d
  mcm_timestep = 0
  ! mcm_getinput.f90:112-127, the reference's own routines
  call mcm_allocate_memory
  call mcm_initialise_database
  ! mcm_initialise_database leaves the cap model's capa and ptype unset, and
  ! the model's first update reads capa (mcm_f3dm25.f90:162); a freshly
  ! mapped allocation holds zeros, which is set here explicitly
  par(:)%capa = 0.0_d
  par(:)%ptype = 0
  call mcm_matin
  call mcm_nodein
  if (mcm_thnode .gt. 0) call mcm_nodehist
  if (mcm_num_transducer .gt. 0) call mcm_transducers_in
  call mcm_inputvel
  if (mcm_boundary) call mcm_inp_symmetry_v2
  if (mcm_contacttype .gt. 0) call mcm_contact_in
  if (mcm_baseaccel) call mcm_baseaccel_in
  close(101)
  mcm_openfiles(1) = 0
 end subroutine probe_read_text_deck

 ! Dump layout (little-endian stream):
 !   int32  magic 'MCMD', step, stage, np, ngp, nummat
 !   f64    dt, dtold, ptime, critts, hmax, kinetice, internale, totale
 !   f64    coord_maxmin(2,3), gridmin(3), gridsize(3)
 !   int32  gridlim(3)
 !   per material: f64 ke, ie, mom(3)
 !   f64 [np][NF] particle fields (row per particle, NF=86)
 !   int32 nnbr(np), g_nnbr(np), llpointer(np)
 !   int32 neighbour ids, particle by particle (nnbr entries), then ghost ids (g_nnbr entries)
 !   ghosts: int32 par(ngp), mat(ngp); f64 [ngp][8] (x3, v3, mass, h), f64 rho(ngp), hold(ngp)
 !   with a cap-model material: int32 'CAP4', f64 capa(np), int32 ptype(np)
 subroutine probe_dump(od, istep, stage)
  character(len=*) :: od
  integer :: istep, stage, u, i, k, m
  character(len=600) :: fname
  real(kind=real_acc) :: row(86)
  write(fname,'(a,"/dump_",i6.6,"_",i2.2,".bin")') od, istep, stage
  open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
  write(u) 1296256324, istep, stage, mcm_np, mcm_ngp, mcm_nummat
  write(u) mcm_dt, mcm_dtold, mcm_ptime, mcm_critts, mcm_hmax, mcm_kinetice, mcm_internale, mcm_totale
  write(u) mcm_coord_maxmin, mcm_gridmin, mcm_gridsize
  write(u) mcm_gridlim
  do m = 1, mcm_nummat
    write(u) mcm_mat(m)%ke, mcm_mat(m)%ie, mcm_mat(m)%mom
  enddo
  do i = 1, mcm_np
    call state_row(i, row)
    write(u) row
  enddo
  write(u) (par(i)%nnbr, i=1,mcm_np)
  write(u) (par(i)%g_nnbr, i=1,mcm_np)
  write(u) (par(i)%llpointer, i=1,mcm_np)
  do i = 1, mcm_np
    if (par(i)%nnbr.gt.0) write(u) (mcm_nbrlist(k,i), k=1,par(i)%nnbr)
  enddo
  if (mcm_boundary) then
    do i = 1, mcm_np
      if (par(i)%g_nnbr.gt.0) write(u) (mcm_g_nbrlist(k,i), k=1,par(i)%g_nnbr)
    enddo
    write(u) (gpar(i)%par, i=1,mcm_ngp)
    write(u) (gpar(i)%mat, i=1,mcm_ngp)
    do i = 1, mcm_ngp
      write(u) gpar(i)%x, gpar(i)%v, gpar(i)%mass, gpar(i)%h
    enddo
    write(u) (gpar(i)%rho, i=1,mcm_ngp)
    write(u) (gpar(i)%hold, i=1,mcm_ngp)
  endif
  ! decks with the cap model (material model 4): 'CAP4', capa(np), ptype(np)
  if (any(mcm_mat(1:mcm_nummat)%model.eq.4)) then
    write(u) 1128353844
    write(u) (par(i)%capa, i=1,mcm_np)
    write(u) (par(i)%ptype, i=1,mcm_np)
  endif
  close(u)
 end subroutine probe_dump

 ! The state rows alone: the dump header, scalars, grid and material totals
 ! of probe_dump, then the np x 86 particle rows, without neighbour lists or
 ! ghosts (which a host driving the GPU wrappers does not hold).
 subroutine probe_dump_rows(fname, istep, stage)
  character(len=*) :: fname
  integer :: istep, stage, u, i, m
  real(kind=real_acc) :: row(86)
  open(newunit=u, file=fname, access='stream', form='unformatted', status='replace')
  write(u) 1296256338, istep, stage, mcm_np, mcm_ngp, mcm_nummat
  write(u) mcm_dt, mcm_dtold, mcm_ptime, mcm_critts, mcm_hmax, mcm_kinetice, mcm_internale, mcm_totale
  write(u) mcm_coord_maxmin, mcm_gridmin, mcm_gridsize
  write(u) mcm_gridlim
  do m = 1, mcm_nummat
    write(u) mcm_mat(m)%ke, mcm_mat(m)%ie, mcm_mat(m)%mom
  enddo
  do i = 1, mcm_np
    call state_row(i, row)
    write(u) row
  enddo
  close(u)
 end subroutine probe_dump_rows

 subroutine state_row(i, row)
  integer :: i
  real(kind=real_acc) :: row(86)
  row(1:3) = par(i)%x
  row(4:6) = par(i)%v
  row(7:9) = par(i)%a
  row(10) = par(i)%rho;    row(11) = par(i)%rhoold
  row(12) = par(i)%h;      row(13) = par(i)%hold
  row(14) = par(i)%p;      row(15) = par(i)%e
  row(16) = par(i)%etry;   row(17) = par(i)%einc
  row(18) = par(i)%c;      row(19) = par(i)%vabs
  row(20) = par(i)%tracerod
  row(21) = par(i)%pcut;   row(22) = par(i)%p_cut
  row(23) = par(i)%efps;   row(24) = par(i)%fail
  row(25) = par(i)%mindist; row(26) = par(i)%critts
  row(27) = par(i)%mass;   row(28) = par(i)%rho0
  row(29:37) = reshape(par(i)%sigma, (/9/))
  row(38:46) = reshape(par(i)%s, (/9/))
  row(47:55) = reshape(par(i)%alfa, (/9/))
  row(56:64) = reshape(par(i)%q, (/9/))
  row(65:73) = reshape(par(i)%rod, (/9/))
  row(74:82) = reshape(par(i)%spin, (/9/))
  row(83) = par(i)%h0
  row(84:86) = par(i)%xzero
 end subroutine state_row

end module ref_deck
