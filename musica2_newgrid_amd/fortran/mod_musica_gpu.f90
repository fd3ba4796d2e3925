!===============================================================================
module mod_musica_gpu
!===============================================================================
  !> Fortran binding (iso_c_binding) of libmusica_gpu.so -- the C-ABI declared
  !> in include/musica_gpu.h.  A MUSICAA host uses it in place of the
  !> runge_kutta call of solver (src/Compute_flow/solver.f90:415-417):
  !>   rc = mgpu_step(h, 1)
  !> after handing over the module arrays rho..rhoe (src/Compute_flow/mod_flow.f90)
  !> and the metrics of grid_finalize (src/Grid/grid_finalize.f90).
  !> Errors are reported through mgpu_check, the analogue of
  !> mpistop(msg,1) (src/Utils/warnstop.f90:20-90).
!===============================================================================
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: MGPU_FACE_PERIODIC = 0, MGPU_FACE_INTERFACE = 1, MGPU_FACE_WALL = 2, &
                                   MGPU_FACE_CHARAC = 3, MGPU_FACE_INLET = 4, MGPU_FACE_SLIP = 5
  integer(c_int32_t), parameter :: MGPU_FILTER_NONE = 0, MGPU_FILTER_SF = 1, MGPU_FILTER_SHOCK = 2
  integer(c_int32_t), parameter :: MGPU_FILTER_ARTVISC = 3, MGPU_FILTER_ARTVISC_SHOCK = 4
  integer(c_int32_t), parameter :: MGPU_PREC_FP64 = 0, MGPU_PREC_MIXED = 1
  integer(c_int32_t), parameter :: MGPU_GEOM_CART = 0, MGPU_GEOM_C3 = 1, MGPU_GEOM_C2 = 2
  integer(c_int32_t), parameter :: MGPU_LAUNCH_EAGER = 1
  integer, parameter :: MGPU_NGH = 5

  type, bind(C) :: mgpu_block_t
     integer(c_int32_t) :: nx, ny, nz
     integer(c_int32_t) :: face_kind(6), nbr_block(6), nbr_face(6), orient(6)
  end type mgpu_block_t

  type, bind(C) :: mgpu_case_t
     integer(c_int32_t) :: nblocks, curvilinear, iorder_visc, filter, is_ducros, nrk, precision, launch_flags
     type(c_ptr) :: blocks
     real(c_double) :: crk(8), deltat, gam, rg, c0, S0, diffscale, cok_ratio
     real(c_double) :: dissip_coeff, dissip_shock, csens
     integer(c_int32_t) :: is_adiab = 1, is_sw_edoh = 0
     real(c_double) :: t_wall = 0.0_c_double
     real(c_double) :: p_ref = 0.0_c_double
     integer(c_int32_t) :: is_wall2 = 0, reserved0 = 0
  end type mgpu_case_t

  ! Hot-path fields of param.ini (read_param, src/File_IO/read_param.f90:88-205)
  type :: param_numerics_t
     integer :: idepart, isolver, nrk, vrk, iirs, nmax, stencil, ibc_scheme, iorder_visc
     real(8) :: timemax, CFL, cpumax, dissip_coeff, dissip_shock, csens
     logical :: is_DRP, is_SF, is_sw_edoh, is_shock, is_ducros
  end type param_numerics_t

  interface
     integer(c_int) function mgpu_create(c, h) bind(C, name='mgpu_create')
       import :: c_int, c_ptr, mgpu_case_t
       type(mgpu_case_t), intent(in) :: c
       type(c_ptr), intent(out) :: h
     end function mgpu_create

     integer(c_int) function mgpu_set_metrics_cart(h, block, idx, idy, idz, idx_v, idy_v, idz_v) &
          bind(C, name='mgpu_set_metrics_cart')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: block
       real(c_double), intent(in) :: idx(*), idy(*), idz(*), idx_v(*), idy_v(*), idz_v(*)
     end function mgpu_set_metrics_cart

     integer(c_int) function mgpu_set_metrics_c3(h, block, m, mv) bind(C, name='mgpu_set_metrics_c3')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: block
       type(c_ptr), intent(in) :: m(10), mv(10)
     end function mgpu_set_metrics_c3

     ! 2.5-D (is_curv): m2 = c_loc of x_ksi, x_eta, y_ksi, y_eta, ijacob (padded 2-D), m2v the *_v set
     integer(c_int) function mgpu_set_metrics_c2(h, block, m2, m2v, idz, idz_v) bind(C, name='mgpu_set_metrics_c2')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: block
       type(c_ptr), intent(in) :: m2(5), m2v(5)
       real(c_double), intent(in) :: idz(*), idz_v(*)
     end function mgpu_set_metrics_c2

     integer(c_int) function mgpu_set_wall(h, block, cextrp2, cextrp3) bind(C, name='mgpu_set_wall')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: block
       real(c_double), intent(in) :: cextrp2(6), cextrp3(6)
     end function mgpu_set_wall

     ! coef: c_loc of the 8 coefficient arrays (nx*nz each) of a c3 j-wall
     integer(c_int) function mgpu_set_wall_c3(h, block, face, coef) bind(C, name='mgpu_set_wall_c3')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: block, face
       type(c_ptr), intent(in) :: coef(8)
     end function mgpu_set_wall_c3

     ! t: c_loc of the 6 dual-tangent arrays (nx*nz each) of a c3 slip j-wall
     integer(c_int) function mgpu_set_wall_c3_slip(h, block, face, t) bind(C, name='mgpu_set_wall_c3_slip')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: block, face
       type(c_ptr), intent(in) :: t(6)
     end function mgpu_set_wall_c3_slip

     integer(c_int) function mgpu_put_state(h, block, rho, rhou, rhov, rhow, rhoe) bind(C, name='mgpu_put_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: block
       real(c_double), intent(in) :: rho(*), rhou(*), rhov(*), rhow(*), rhoe(*)
     end function mgpu_put_state

     integer(c_int) function mgpu_get_state(h, block, rho, rhou, rhov, rhow, rhoe) bind(C, name='mgpu_get_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: block
       real(c_double), intent(out) :: rho(*), rhou(*), rhov(*), rhow(*), rhoe(*)
     end function mgpu_get_state

     integer(c_int) function mgpu_step(h, nsteps) bind(C, name='mgpu_step')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: nsteps
     end function mgpu_step

     integer(c_int) function mgpu_autotune(h, flags) bind(C, name='mgpu_autotune')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: flags
     end function mgpu_autotune

     integer(c_int) function mgpu_sync(h) bind(C, name='mgpu_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mgpu_sync

     type(c_ptr) function mgpu_last_error(h) bind(C, name='mgpu_last_error')
       import :: c_ptr
       type(c_ptr), value :: h
     end function mgpu_last_error

     subroutine mgpu_destroy(h) bind(C, name='mgpu_destroy')
       import :: c_ptr
       type(c_ptr), value :: h
     end subroutine mgpu_destroy

     integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  !=============================================================================
  subroutine mgpu_check(rc, h, what)
  !=============================================================================
    !> Stop with the library's message when rc /= 0 (mpistop analogue).
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    character(len=512) :: msg
    integer :: n, i

    if (rc == 0) return
    p = mgpu_last_error(h)
    msg = ''
    if (c_associated(p)) then
       n = int(min(c_strlen(p), 512_c_size_t))
       call c_f_pointer(p, s, [n])
       do i = 1, n
          msg(i:i) = s(i)
       enddo
    endif
    write(*,'(a,a,a,i0,a,a)') 'mgpu error in ', trim(what), ' (rc=', rc, '): ', trim(msg)
    error stop 1
  end subroutine mgpu_check

  !=============================================================================
  subroutine read_param_numerics(fname, p)
  !=============================================================================
    !> Read the solver/numerics block of the reference's param.ini with the
    !> same positional record sequence as read_param
    !> (src/File_IO/read_param.f90:88-205): every comment line is consumed
    !> by one record, every data line is list-directed.
    character(len=*), intent(in) :: fname
    type(param_numerics_t), intent(out) :: p
    integer :: u, ios
    real(8) :: th(5)
    logical :: li(3), is_dtvar, is_dtlocal, is_adjoint_block, verbose
    integer :: neval_dt
    real(8) :: c_dtloc

    open(newunit=u, file=fname, status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(*,*) 'param file not found: ', trim(fname)
       error stop 1
    endif
    call skip(u, 29)                       ! header and mode comments (:89-117)
    read(u,*) p%idepart                    ! :118
    call skip(u, 1); read(u,*) p%isolver
    call skip(u, 1); read(u,*) is_adjoint_block
    call skip(u, 1); read(u,*) verbose
    call skip(u, 8); read(u,*) p%nrk, p%vrk
    call skip(u, 1); read(u,*) p%iirs
    call skip(u, 1); read(u,*) th
    call skip(u, 1); read(u,*) li
    call skip(u, 4); read(u,*) p%nmax, p%timemax
    call skip(u, 1); read(u,*) is_dtvar, neval_dt
    call skip(u, 1); read(u,*) is_dtlocal, c_dtloc
    call skip(u, 1); read(u,*) p%CFL
    call skip(u, 1); read(u,*) p%cpumax
    call skip(u, 4); read(u,*) p%stencil, p%is_DRP
    call skip(u, 1); read(u,*) p%ibc_scheme
    call skip(u, 1); read(u,*) p%iorder_visc
    call skip(u, 1); read(u,*) p%is_SF
    call skip(u, 1); read(u,*) p%is_sw_edoh
    call skip(u, 2); read(u,*) p%dissip_coeff
    call skip(u, 1); read(u,*) p%is_shock, p%dissip_shock
    call skip(u, 1); read(u,*) p%is_ducros, p%csens
    close(u)
    ! protection against divide by zero, as read_param does (:205)
    p%csens = max(p%csens, 0.01d0)
  end subroutine read_param_numerics

  subroutine skip(u, n)
    integer, intent(in) :: u, n
    integer :: i
    do i = 1, n
       read(u,*)
    enddo
  end subroutine skip

end module mod_musica_gpu
