!===============================================================================
program musica_gpu_tgv
!===============================================================================
  !> Fortran host of the GPU time advance on the Taylor-Green vortex
  !> (config: periodic Cartesian cube, Mach 0.1, Re 1600, T_ref 300 K).
  !> Reads the numerics of a MUSICAA param.ini, builds the case the way the
  !> reference does (reference quantities setupref.f90:39-69,258-266,357; IC
  !> add_tgv mod_init_flow.f90:577-620; metrics grid_finalize.f90:234-238;
  !> deltat timestep_new.f90:256-288), then runs the solver loop with
  !> mgpu_step in place of runge_kutta.
  !>   usage: musica_gpu_tgv param.ini n nsteps out_prefix
  !> Writes out_prefix.init.bin / .final.bin (5 padded arrays each, stream)
  !> and out_prefix.case.txt (the scalars of the run).
!===============================================================================
  use iso_c_binding
  use mod_musica_gpu
  implicit none
  integer, parameter :: ngh = MGPU_NGH
  real(8), parameter :: pi = 3.141592653589793238462643383279502884d0
  type(param_numerics_t) :: prm
  type(mgpu_case_t) :: cs
  type(mgpu_block_t), target :: blk(1)
  type(c_ptr) :: h
  character(len=256) :: pfile, arg, prefix
  integer :: n, nsteps, i, j, k, u, rc, l
  real(8), allocatable, target :: rho(:,:,:), rhou(:,:,:), rhov(:,:,:), rhow(:,:,:), rhoe(:,:,:)
  real(8), allocatable :: idx(:)
  real(8) :: T_ref, rho_ref, Mach, Re, gam, rg, Pr, p_ref, c_ref, u_ref, mu_ref, muw_ref
  real(8) :: c0, S0, dx, x, y, z, uu, vv, ww, pp, TT, rr, umax, dt, ek, mass
  real(8) :: t0, t1, g(6)

  call get_command_argument(1, pfile)
  call get_command_argument(2, arg); read(arg,*) n
  call get_command_argument(3, arg); read(arg,*) nsteps
  call get_command_argument(4, prefix)
  call read_param_numerics(trim(pfile), prm)
  if (nsteps <= 0) nsteps = prm%nmax

  ! reference quantities (air, feos_air.ini; Sutherland smod_tranprop_init.f90:33-40)
  T_ref = 300.0d0; rho_ref = 1.177d0; Mach = 0.1d0; Re = 1600.0d0
  gam = 1.4d0; rg = 287.0d0; Pr = 0.71d0
  c0 = 1.4579326545176254d-06; S0 = 110.4d0
  p_ref = rho_ref*rg*T_ref
  c_ref = sqrt(gam*rg*T_ref)
  u_ref = Mach*c_ref
  mu_ref = rho_ref*u_ref*1.0d0/Re
  muw_ref = c0*sqrt(T_ref**3)/(T_ref+S0)

  allocate(rho(1-ngh:n+ngh,1-ngh:n+ngh,1-ngh:n+ngh))
  allocate(rhou, rhov, rhow, rhoe, mold=rho)
  rho = 0; rhou = 0; rhov = 0; rhow = 0; rhoe = 0
  dx = 2.0d0*pi/n
  umax = 0.0d0
  do k = 1, n
     z = (k-1)*dx
     do j = 1, n
        y = (j-1)*dx
        do i = 1, n
           x = (i-1)*dx
           uu =  u_ref*sin(x)*cos(y)*cos(z)
           vv = -u_ref*cos(x)*sin(y)*cos(z)
           ww = 0.0d0
           pp = p_ref + rho_ref*u_ref**2/16.0d0*(cos(2.0d0*z)+2.0d0)*(cos(2.0d0*x)+cos(2.0d0*y))
           TT = pp/p_ref*T_ref
           rr = pp/(rg*TT)
           rho(i,j,k) = rr
           rhou(i,j,k) = rr*uu; rhov(i,j,k) = rr*vv; rhow(i,j,k) = rr*ww
           rhoe(i,j,k) = pp/(gam-1.0d0) + 0.5d0*rr*(uu*uu+vv*vv+ww*ww)
           umax = max(umax, sqrt(uu*uu+vv*vv+ww*ww) + sqrt(gam*pp/rr))
        enddo
     enddo
  enddo
  dt = prm%CFL*dx/umax

  allocate(idx(1-ngh:n+ngh))
  idx = 1.0d0/dx

  ! case description (flags from param.ini, assign_procedures.f90:161-219)
  blk(1)%nx = n; blk(1)%ny = n; blk(1)%nz = n
  blk(1)%face_kind = MGPU_FACE_PERIODIC
  blk(1)%nbr_block = 0
  blk(1)%nbr_face = [1, 0, 3, 2, 5, 4]
  blk(1)%orient = 0
  cs%nblocks = 1
  cs%curvilinear = 0
  cs%iorder_visc = prm%iorder_visc
  ! num_dissip as assign_procedures binds it (is_dissip_in_increments = F):
  ! filtering_11pts, filter_o10_shock, or artvisc_o9 / artvisc_o9_shock when is_SF = F
  if (prm%is_SF) then
     cs%filter = MGPU_FILTER_SF
     if (prm%is_shock) cs%filter = MGPU_FILTER_SHOCK
  else
     cs%filter = MGPU_FILTER_ARTVISC
     if (prm%is_shock) cs%filter = MGPU_FILTER_ARTVISC_SHOCK
  endif
  cs%is_ducros = merge(1, 0, prm%is_ducros)
  cs%is_sw_edoh = merge(1, 0, prm%is_sw_edoh)
  cs%nrk = prm%nrk
  cs%precision = MGPU_PREC_FP64
  cs%launch_flags = 0
  cs%blocks = c_loc(blk(1))
  cs%crk = 0.0d0
  select case (prm%nrk)
  case (4)   ! init_RK, mod_time.f90:165-169
     cs%crk(1:4) = [1.0d0/4.0d0, 1.0d0/3.0d0, 1.0d0/2.0d0, 1.0d0]
  case (6)   ! DRP scheme, mod_time.f90:253-294: amplification polynomial sum g_m z^m
     g = [1.0d0, 0.5d0, 0.165919771368d0, 0.040919732041d0, 0.007555704391d0, 0.000891421261d0]
     cs%crk(6) = g(1)
     cs%crk(5) = g(2)
     cs%crk(4) = g(3)/cs%crk(5)
     cs%crk(3) = g(4)/(cs%crk(4)*cs%crk(5))
     cs%crk(2) = g(5)/(cs%crk(4)*cs%crk(3)*cs%crk(5))
     cs%crk(1) = g(6)/(cs%crk(4)*cs%crk(3)*cs%crk(5)*cs%crk(2))
  case default
     write(*,*) 'nrk must be 4 or 6 (init_RK, mod_time.f90:121-331)'
     error stop 1
  end select
  cs%deltat = dt
  cs%gam = gam; cs%rg = rg; cs%c0 = c0; cs%S0 = S0
  cs%diffscale = mu_ref/muw_ref
  cs%cok_ratio = gam*rg/(gam-1.0d0)/Pr
  cs%dissip_coeff = prm%dissip_coeff
  cs%dissip_shock = prm%dissip_shock
  cs%csens = prm%csens

  rc = mgpu_create(cs, h)
  call mgpu_check(rc, c_null_ptr, 'mgpu_create')
  rc = mgpu_set_metrics_cart(h, 0, idx, idx, idx, idx, idx, idx)
  call mgpu_check(rc, h, 'mgpu_set_metrics_cart')
  rc = mgpu_put_state(h, 0, rho, rhou, rhov, rhow, rhoe)
  call mgpu_check(rc, h, 'mgpu_put_state')
  rc = mgpu_get_state(h, 0, rho, rhou, rhov, rhow, rhoe)   ! ghosts as filled on the device
  call mgpu_check(rc, h, 'mgpu_get_state')
  call dump(trim(prefix)//'.init.bin')

  call cpu_time(t0)
  ! solver loop (solver.f90:394-535) with the step on the GPU
  do l = 1, nsteps
     rc = mgpu_step(h, 1)
     call mgpu_check(rc, h, 'mgpu_step')
  enddo
  rc = mgpu_sync(h)
  call mgpu_check(rc, h, 'mgpu_sync')
  call cpu_time(t1)
  rc = mgpu_get_state(h, 0, rho, rhou, rhov, rhow, rhoe)
  call mgpu_check(rc, h, 'mgpu_get_state')
  call mgpu_destroy(h)
  call dump(trim(prefix)//'.final.bin')

  ! TGV statistics (stats_tgv, src/File_IO/stats.f90:37-112): mass and kinetic energy
  mass = sum(rho(1:n,1:n,1:n))/n**3
  ek = 0.5d0*sum((rhou(1:n,1:n,1:n)**2+rhov(1:n,1:n,1:n)**2+rhow(1:n,1:n,1:n)**2)/rho(1:n,1:n,1:n))/n**3
  write(*,'(a,i0,a,i0,a,es12.5,a,es22.15,a,es22.15)') 'TGV n=', n, ' steps=', nsteps, ' dt=', dt, &
       ' mean rho=', mass, ' mean Ek=', ek
  write(*,'(a,f10.3,a)') 'host time in the step loop: ', (t1-t0)*1.0d3, ' ms'

  open(newunit=u, file=trim(prefix)//'.case.txt', status='replace')
  write(u,'(a,i0)') 'n ', n
  write(u,'(a,i0)') 'nsteps ', nsteps
  write(u,'(a,es25.17)') 'dt ', dt
  write(u,'(a,es25.17)') 'diffscale ', cs%diffscale
  write(u,'(a,i0)') 'iorder_visc ', cs%iorder_visc
  write(u,'(a,i0)') 'filter ', cs%filter
  write(u,'(a,es25.17)') 'chi ', cs%dissip_coeff
  write(u,'(a,es25.17)') 'dissip_shock ', cs%dissip_shock
  write(u,'(a,es25.17)') 'csens ', cs%csens
  write(u,'(a,i0)') 'ducros ', cs%is_ducros
  close(u)

contains

  subroutine dump(fname)
    character(len=*), intent(in) :: fname
    integer :: w
    open(newunit=w, file=fname, access='stream', form='unformatted', status='replace')
    write(w) rho, rhou, rhov, rhow, rhoe
    close(w)
  end subroutine dump

end program musica_gpu_tgv
