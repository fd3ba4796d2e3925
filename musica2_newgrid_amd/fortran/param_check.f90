!===============================================================================
program param_check
!===============================================================================
  !> Print the numerics block of a param.ini as parsed by read_param_numerics
  !> (one "key value" per line), so that the positional reader is testable
  !> without a GPU.   usage: param_check param.ini
!===============================================================================
  use mod_musica_gpu
  implicit none
  type(param_numerics_t) :: p
  character(len=256) :: f
  call get_command_argument(1, f)
  call read_param_numerics(trim(f), p)
  write(*,'(a,i0)') 'idepart ', p%idepart
  write(*,'(a,i0)') 'isolver ', p%isolver
  write(*,'(a,i0,1x,i0)') 'nrk ', p%nrk, p%vrk
  write(*,'(a,i0)') 'iirs ', p%iirs
  write(*,'(a,i0)') 'nmax ', p%nmax
  write(*,'(a,es23.15)') 'CFL ', p%CFL
  write(*,'(a,i0)') 'stencil ', p%stencil
  write(*,'(a,l1)') 'is_DRP ', p%is_DRP
  write(*,'(a,i0)') 'iorder_visc ', p%iorder_visc
  write(*,'(a,l1)') 'is_SF ', p%is_SF
  write(*,'(a,es23.15)') 'dissip_coeff ', p%dissip_coeff
  write(*,'(a,l1)') 'is_shock ', p%is_shock
  write(*,'(a,es23.15)') 'dissip_shock ', p%dissip_shock
  write(*,'(a,l1)') 'is_ducros ', p%is_ducros
  write(*,'(a,es23.15)') 'csens ', p%csens
end program param_check
