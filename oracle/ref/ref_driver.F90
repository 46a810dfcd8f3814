! Golden-vector driver for the reference tracer advection (test infrastructure only).
!
! Reads a case bundle written by oracle/ref/bundle.py (dims.txt + raw little-endian arrays in
! Fortran order), fills the reference's T_MESH / T_PARTIT / T_TRACER / T_DYN from it verbatim and
! advances every tracer nsteps times with the reference kernels:
!
!   per tracer and step (caller semantics of src/fesom_analytic.F90:404-437):
!     del_ttf_advhoriz = del_ttf_advvert = 0
!     valuesAB by ab_mode: 0  valuesAB := values
!                          1  valuesAB frozen at the initial values (the shipped drivers,
!                             src/fesom_analytic.F90:335)
!                          2  FESOM2's AB2 interpolation valuesAB = -(0.5+epsilon)*valuesold
!                             + (1.5+epsilon)*values, valuesold := values (epsilon of
!                             lib/oce_modules.F90:75; the dwarf never computes it)
!     tr_xy    := P1 gradient of valuesAB with gradient_sca (FESOM2's tracer_gradient_elements)
!     call fill_up_dn_grad                        (lib/oce_muscl_adv.F90:361-530)
!     call do_oce_adv_tra(dt, uv, w, wi, we, ...) (lib/oce_adv_tra_driver.F90:23-294), with
!                                                 dynamics%use_wsplit from the bundle
!     values += (del_ttf_advhoriz + del_ttf_advvert) / hnode_new
!
! Usage: ref_driver <bundle-dir> [<restart-dir>]
! Outputs (in the bundle directory):
!   out_values_sNNNN.bin   all tracers (nl-1, nod2D, ntr) after each requested step
!   out_w_*.bin            work arrays of tracer 1 at step 1 (edge_up_dn_grad before advection,
!                          fct_LO, limited adv_flux_hor / adv_flux_ver, del_ttf_advhoriz/advvert)
program ref_driver
  use mpi
  use o_PARAM
  use o_ARRAYS, only: tr_xy
  use MOD_MESH
  use MOD_PARTIT
  use MOD_PARSUP
  use MOD_TRACER
  use MOD_DYN
  use par_support_interfaces, only: par_init
  use oce_adv_tra_driver_module, only: do_oce_adv_tra
  use oce_adv_tra_fct_module, only: oce_adv_tra_fct_init
  use oce_muscl_adv_module, only: fill_up_dn_grad
  use restart_derivedtype_module, only: write_all_bin_restarts, read_all_bin_restarts
  use fortran_utils, only: int_to_txt, mkdir
#if defined(_OPENMP)
  use omp_lib
#endif
  implicit none

  type(t_mesh),   target :: mesh
  type(t_partit), target :: partit
  type(t_tracer), target :: tracers
  type(t_dyn),    target :: dyn
  character(len=4096)    :: dir, rdir, arg
  logical                :: rmode
  integer, dimension(3)  :: ctime = (/0, 0, 0/)
  integer :: ierr, nl, nod2D, elem2D, edge2D, max_adj, ntr, nsteps, nsave, i, n, k, e, istep, nz
  integer :: save_work, timing, use_wsplit, ab_mode
  real(kind=WP), allocatable :: vold(:,:,:)
  integer, allocatable :: save_steps(:)
  real(kind=WP) :: dt
  character(len=20) :: hor, ver, lim
  real(kind=WP) :: ph, pv
  integer :: u
  real(kind=WP), allocatable :: tbuf(:,:,:)
  double precision :: t0, t_grad, t_adv, t_upd
  integer(8) :: nlev_total

  call MPI_Init(ierr)
  call par_init(partit)
  call get_command_argument(1, dir)
  if (len_trim(dir) == 0) then
     write(*,*) 'usage: ref_driver <bundle-dir> [<restart-dir>] | --restart <restart-dir> <out-dir> <nsteps> <dt>'
     call MPI_Abort(MPI_COMM_WORLD, 2, ierr)
  end if
  rmode = trim(dir) == '--restart'
  if (rmode) then
     ! state from derived-type restart files; values after the last step into <out-dir>
     call get_command_argument(2, rdir)
     call get_command_argument(3, dir)
     call get_command_argument(4, arg)
     read(arg, *) nsteps
     call get_command_argument(5, arg)
     read(arg, *) dt
     call read_all_bin_restarts(trim(rdir)//'/np'//int_to_txt(partit%npes), partit=partit, mesh=mesh, &
          dynamics=dyn, tracers=tracers)
     nl = mesh%nl; nod2D = mesh%nod2D; elem2D = mesh%elem2D; edge2D = mesh%edge2D
     ntr = tracers%num_tracers
     save_work = 0; timing = 0; ab_mode = 0
     nsave = 1
     allocate(save_steps(1))
     save_steps(1) = nsteps
     rdir = ''
     if (.not. allocated(mesh%zbar_n_bot)) then
        allocate(mesh%zbar_n_bot(nod2D))
        do k = 1, nod2D
           mesh%zbar_n_bot(k) = mesh%zbar_3d_n(mesh%nlevels_nod2D(k), k)
        end do
     end if
     allocate(tbuf(nl-1, nod2D, ntr), vold(nl-1, nod2D, ntr))
     do n = 1, ntr
        vold(:, :, n) = tracers%data(n)%values
        if (.not. allocated(tracers%data(n)%valuesAB)) allocate(tracers%data(n)%valuesAB(nl-1, nod2D))
     end do
     if (allocated(tracers%work%fct_LO)) deallocate(tracers%work%fct_LO, tracers%work%adv_flux_hor, &
          tracers%work%adv_flux_ver, tracers%work%fct_ttf_max, tracers%work%fct_ttf_min, &
          tracers%work%fct_plus, tracers%work%fct_minus)
  else
  open(newunit=u, file=trim(dir)//'/dims.txt', status='old', action='read')
  read(u,*) nl, nod2D, elem2D, edge2D, max_adj
  read(u,*) ntr, nsteps, dt
  read(u,*) save_work, timing, use_wsplit, ab_mode
  read(u,*) nsave
  allocate(save_steps(nsave))
  if (nsave > 0) read(u,*) save_steps
  tracers%num_tracers = ntr
  allocate(tracers%data(ntr))
  do n = 1, ntr
     read(u,*) hor, ver, lim, ph, pv
     tracers%data(n)%tra_adv_hor = hor
     tracers%data(n)%tra_adv_ver = ver
     tracers%data(n)%tra_adv_lim = lim
     tracers%data(n)%tra_adv_ph  = ph
     tracers%data(n)%tra_adv_pv  = pv
  end do
  close(u)

  mesh%nl = nl; mesh%nod2D = nod2D; mesh%elem2D = elem2D; mesh%edge2D = edge2D
  mesh%edge2D_in = edge2D
  partit%myDim_nod2D = nod2D;   partit%eDim_nod2D = 0
  partit%myDim_elem2D = elem2D; partit%eDim_elem2D = 0; partit%eXDim_elem2D = 0
  partit%myDim_edge2D = edge2D; partit%eDim_edge2D = 0

  allocate(mesh%elem2D_nodes(3, elem2D), mesh%edges(2, edge2D), mesh%edge_tri(2, edge2D))
  allocate(mesh%nod_in_elem2D(max_adj, nod2D), mesh%nod_in_elem2D_num(nod2D))
  allocate(mesh%nlevels(elem2D), mesh%ulevels(elem2D))
  allocate(mesh%nlevels_nod2D(nod2D), mesh%ulevels_nod2D(nod2D))
  allocate(mesh%nlevels_nod2D_min(nod2D), mesh%ulevels_nod2D_max(nod2D))
  allocate(mesh%elem_area(elem2D), mesh%area(nl, nod2D), mesh%areasvol(nl, nod2D))
  allocate(mesh%edge_dxdy(2, edge2D), mesh%edge_cross_dxdy(4, edge2D))
  allocate(mesh%elem_cos(elem2D), mesh%gradient_sca(6, elem2D))
  allocate(mesh%hnode(nl-1, nod2D), mesh%hnode_new(nl-1, nod2D), mesh%helem(nl-1, elem2D))
  allocate(mesh%zbar_3d_n(nl, nod2D), mesh%Z_3d_n(nl-1, nod2D))
  call rd_i2(mesh%elem2D_nodes, 'elem2D_nodes')
  call rd_i2(mesh%edges, 'edges')
  call rd_i2(mesh%edge_tri, 'edge_tri')
  call rd_i2(mesh%nod_in_elem2D, 'nod_in_elem2D')
  call rd_i1(mesh%nod_in_elem2D_num, 'nod_in_elem2D_num')
  call rd_i1(mesh%nlevels, 'nlevels')
  call rd_i1(mesh%ulevels, 'ulevels')
  call rd_i1(mesh%nlevels_nod2D, 'nlevels_nod2D')
  call rd_i1(mesh%ulevels_nod2D, 'ulevels_nod2D')
  call rd_i1(mesh%nlevels_nod2D_min, 'nlevels_nod2D_min')
  call rd_i1(mesh%ulevels_nod2D_max, 'ulevels_nod2D_max')
  call rd_r1(mesh%elem_area, 'elem_area')
  call rd_r2(mesh%area, 'area')
  call rd_r2(mesh%areasvol, 'areasvol')
  call rd_r2(mesh%edge_dxdy, 'edge_dxdy')
  call rd_r2(mesh%edge_cross_dxdy, 'edge_cross_dxdy')
  call rd_r1(mesh%elem_cos, 'elem_cos')
  call rd_r2(mesh%gradient_sca, 'gradient_sca')
  call rd_r2(mesh%hnode, 'hnode')
  call rd_r2(mesh%hnode_new, 'hnode_new')
  call rd_r2(mesh%helem, 'helem')
  call rd_r2(mesh%zbar_3d_n, 'zbar_3d_n')
  call rd_r2(mesh%Z_3d_n, 'Z_3d_n')
  ! bottom depth per node: adv_tra_vert_impl reads it (into arrays its solve never uses)
  allocate(mesh%zbar_n_bot(nod2D))
  do k = 1, nod2D
     mesh%zbar_n_bot(k) = mesh%zbar_3d_n(mesh%nlevels_nod2D(k), k)
  end do

  allocate(dyn%uv(2, nl-1, elem2D), dyn%w(nl, nod2D), dyn%w_i(nl, nod2D), dyn%w_e(nl, nod2D))
  call rd_r3(dyn%uv, 'uv')
  call rd_r2(dyn%w, 'w')
  call rd_r2(dyn%w_i, 'wi')
  call rd_r2(dyn%w_e, 'we')
  dyn%use_wsplit = (use_wsplit /= 0)

  allocate(tbuf(nl-1, nod2D, ntr))
  call rd_r3(tbuf, 'values')
  do n = 1, ntr
     allocate(tracers%data(n)%values(nl-1, nod2D), tracers%data(n)%valuesAB(nl-1, nod2D))
     tracers%data(n)%values = tbuf(:, :, n)
     tracers%data(n)%valuesAB = tbuf(:, :, n)
  end do
  allocate(vold(nl-1, nod2D, ntr))
  vold = tbuf
  allocate(tracers%work%edge_up_dn_tri(2, edge2D), tracers%work%nboundary_lay(nod2D))
  call rd_i2(tracers%work%edge_up_dn_tri, 'edge_up_dn_tri')
  call rd_i1(tracers%work%nboundary_lay, 'nboundary_lay')
  end if
#if defined(_OPENMP)
  ! the node locks of the reference's OpenMP scatter (as lib/gen_modules_partitioning.F90:578-585)
  allocate(partit%plock(max(nod2D, elem2D)))
  do i = 1, size(partit%plock)
     call omp_init_lock(partit%plock(i))
  end do
#endif

  if (.not. allocated(tracers%work%del_ttf_advhoriz)) &
       allocate(tracers%work%del_ttf_advhoriz(nl-1, nod2D), tracers%work%del_ttf_advvert(nl-1, nod2D))
  call oce_adv_tra_fct_init(tracers%work, partit, mesh)
  if (.not. allocated(tracers%work%edge_up_dn_grad)) allocate(tracers%work%edge_up_dn_grad(4, nl-1, edge2D))
  tracers%work%edge_up_dn_grad = 0.0_MP
  allocate(tr_xy(2, nl-1, elem2D))
  tr_xy = 0.0_WP

  t_grad = 0.d0; t_adv = 0.d0; t_upd = 0.d0
  do istep = 1, nsteps
     do n = 1, ntr
        tracers%work%del_ttf_advhoriz = 0.0_MP
        tracers%work%del_ttf_advvert  = 0.0_MP
        t0 = MPI_Wtime()
        if (ab_mode == 0) then
           tracers%data(n)%valuesAB = tracers%data(n)%values
        else if (ab_mode == 2) then
           tracers%data(n)%valuesAB = -(0.5_WP+epsilon)*vold(:, :, n) + (1.5_WP+epsilon)*tracers%data(n)%values
           vold(:, :, n) = tracers%data(n)%values
        end if
        call gradients(tracers%data(n)%valuesAB)
        call fill_up_dn_grad(tracers%work, partit, mesh)
        if (save_work /= 0 .and. istep == 1 .and. n == 1) &
             call wr_r3(tracers%work%edge_up_dn_grad, 'out_w_edge_up_dn_grad')
        t_grad = t_grad + (MPI_Wtime() - t0)
        t0 = MPI_Wtime()
        call do_oce_adv_tra(dt, dyn%uv, dyn%w, dyn%w_i, dyn%w_e, n, dyn, tracers, partit, mesh)
        t_adv = t_adv + (MPI_Wtime() - t0)
        if (save_work /= 0 .and. istep == 1 .and. n == 1) then
           call wr_r2(tracers%work%fct_LO, 'out_w_fct_LO')
           call wr_r2(tracers%work%adv_flux_hor, 'out_w_adv_flux_hor')
           call wr_r2(tracers%work%adv_flux_ver, 'out_w_adv_flux_ver')
           call wr_r2(tracers%work%del_ttf_advhoriz, 'out_w_dttf_h')
           call wr_r2(tracers%work%del_ttf_advvert, 'out_w_dttf_v')
        end if
        t0 = MPI_Wtime()
        do k = 1, nod2D
           do nz = mesh%ulevels_nod2D(k), mesh%nlevels_nod2D(k) - 1
              tracers%data(n)%values(nz, k) = tracers%data(n)%values(nz, k) + &
                   (tracers%work%del_ttf_advhoriz(nz, k) + tracers%work%del_ttf_advvert(nz, k)) / &
                   mesh%hnode_new(nz, k)
           end do
        end do
        t_upd = t_upd + (MPI_Wtime() - t0)
     end do
     do i = 1, nsave
        if (save_steps(i) == istep) then
           do n = 1, ntr
              tbuf(:, :, n) = tracers%data(n)%values
           end do
           call wr_r3(tbuf, 'out_values_s'//step_tag(istep))
        end if
     end do
  end do

  if (timing /= 0) then
     nlev_total = 0
     do k = 1, nod2D
        nlev_total = nlev_total + (mesh%nlevels_nod2D(k) - mesh%ulevels_nod2D(k))
     end do
     open(newunit=u, file=trim(dir)//'/out_timing.txt', status='replace', action='write')
     write(u, '(A,I0)') 'node_levels_per_tracer ', nlev_total
     write(u, '(A,I0)') 'tracer_steps ', nsteps * ntr
     write(u, '(A,ES16.8)') 'seconds_gradients ', t_grad
     write(u, '(A,ES16.8)') 'seconds_advection ', t_adv
     write(u, '(A,ES16.8)') 'seconds_update ', t_upd
     close(u)
  end if
  ! optional second argument: write the state as the reference's derived-type restart files
  ! (lib/io_restart_derivedtype.F90) into <dir>/np<npes>/ and <dir>/meta.time
  if (.not. rmode) call get_command_argument(2, rdir)
  if (len_trim(rdir) > 0) then
     call mkdir(trim(rdir))
     call mkdir(trim(rdir)//'/np'//int_to_txt(partit%npes))
     call write_all_bin_restarts(ctime, trim(rdir)//'/np'//int_to_txt(partit%npes), trim(rdir)//'/meta.time', &
          partit=partit, mesh=mesh, dynamics=dyn, tracers=tracers)
  end if
  call MPI_Finalize(ierr)

contains

  function step_tag(s) result(tag)
    integer, intent(in) :: s
    character(len=4) :: tag
    write(tag, '(I4.4)') s
  end function step_tag

  ! tr_xy(:, nz, elem) over the element's wet levels: sum over its three nodes of
  ! gradient_sca * valuesAB, accumulated in node order.
  subroutine gradients(ttf)
    real(kind=WP), intent(in) :: ttf(:,:)
    integer :: el, lev, en(3)
    do el = 1, elem2D
       en = mesh%elem2D_nodes(:, el)
       do lev = mesh%ulevels(el), mesh%nlevels(el) - 1
          tr_xy(1, lev, el) = mesh%gradient_sca(1, el)*ttf(lev, en(1)) + &
               mesh%gradient_sca(2, el)*ttf(lev, en(2)) + mesh%gradient_sca(3, el)*ttf(lev, en(3))
          tr_xy(2, lev, el) = mesh%gradient_sca(4, el)*ttf(lev, en(1)) + &
               mesh%gradient_sca(5, el)*ttf(lev, en(2)) + mesh%gradient_sca(6, el)*ttf(lev, en(3))
       end do
    end do
  end subroutine gradients

  subroutine rd_i1(a, name)
    integer, intent(out) :: a(:)
    character(len=*), intent(in) :: name
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='old', action='read')
    read(uu) a
    close(uu)
  end subroutine rd_i1
  subroutine rd_i2(a, name)
    integer, intent(out) :: a(:,:)
    character(len=*), intent(in) :: name
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='old', action='read')
    read(uu) a
    close(uu)
  end subroutine rd_i2
  subroutine rd_r1(a, name)
    real(kind=WP), intent(out) :: a(:)
    character(len=*), intent(in) :: name
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='old', action='read')
    read(uu) a
    close(uu)
  end subroutine rd_r1
  subroutine rd_r2(a, name)
    real(kind=WP), intent(out) :: a(:,:)
    character(len=*), intent(in) :: name
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='old', action='read')
    read(uu) a
    close(uu)
  end subroutine rd_r2
  subroutine rd_r3(a, name)
    real(kind=WP), intent(out) :: a(:,:,:)
    character(len=*), intent(in) :: name
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='old', action='read')
    read(uu) a
    close(uu)
  end subroutine rd_r3
  subroutine wr_r2(a, name)
    real(kind=WP), intent(in) :: a(:,:)
    character(len=*), intent(in) :: name
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='replace', action='write')
    write(uu) a
    close(uu)
  end subroutine wr_r2
  subroutine wr_r3(a, name)
    real(kind=WP), intent(in) :: a(:,:,:)
    character(len=*), intent(in) :: name
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', &
         status='replace', action='write')
    write(uu) a
    close(uu)
  end subroutine wr_r3
end program ref_driver
