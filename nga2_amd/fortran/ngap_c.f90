!> iso_c_binding interface to libngapsolve.so (include/ngapsolve.h).
!> Derived types mirror the C structs field by field (bind(C)); functions
!> return the C status code (0 ok, <0 misuse, >0 runtime error).
module ngap_c
   use, intrinsic :: iso_c_binding
   implicit none
   private

   ! methods (hypre_str_class.f90:17-22 in NGA2)
   integer, parameter, public ::       pfmg=1
   integer, parameter, public ::   pcg_pfmg=2
   integer, parameter, public :: gmres_pfmg=3
   integer, parameter, public ::        smg=4
   integer, parameter, public ::    pcg_smg=5
   integer, parameter, public ::  gmres_smg=6

   ! precision modes
   integer, parameter, public :: ngap_fp64=0
   integer, parameter, public :: ngap_mixed=1
   ! correction stage flags (ngap_set_stage)
   integer, parameter, public :: ngap_stage_guess=1
   integer, parameter, public :: ngap_stage_precond=2

   ! status of the last solve
   integer, parameter, public :: ngap_status_converged=0
   integer, parameter, public :: ngap_status_maxit=1
   integer, parameter, public :: ngap_status_stagnated=2

   type, bind(C), public :: ngap_grid_t
      integer(c_int) :: nx,ny,nz
      integer(c_int) :: imin_,imax_,jmin_,jmax_,kmin_,kmax_
      integer(c_int) :: imino_,imaxo_,jmino_,jmaxo_,kmino_,kmaxo_
      integer(c_int) :: xper,yper,zper
   end type ngap_grid_t

   type, bind(C), public :: ngap_opts_t
      integer(c_int) :: precision,nu_pre,nu_post,coarsest_max,check_finite,verbose,graphs,kdim,skip_relax
   end type ngap_opts_t

   type, bind(C), public :: ngap_stats_t
      integer(c_int) :: status,it,nlevels,coarse_singular,precision,fallback,restarts,stage_calls
      integer(c_int) :: skip_relax,relaxed_levels
      real(c_double) :: rerr,true_rerr,aerr,cnorm_ratio,asym,setup_ms,solve_ms,copy_ms
      integer(c_long_long) :: cells,device_bytes
   end type ngap_stats_t

   public :: ngap_default_opts,ngap_create,ngap_init,ngap_setup,ngap_solve,ngap_set_tolerance,ngap_set_stage
   public :: ngap_destroy,ngap_last_stats,ngap_strerror_f
   ! fourier3d on the GPU (libngapfft.so, include/ngapfft.h)
   public :: ngap_fft_create,ngap_fft_setup,ngap_fft_solve,ngap_fft_destroy
   ! hypre_uns on the GPU (libngapuns.so, include/ngapuns.h)
   public :: ngap_uns_default_opts,ngap_uns_create,ngap_uns_init,ngap_uns_setup,ngap_uns_solve,ngap_uns_destroy
   public :: ngap_uns_rows,ngap_uns_index

   type, bind(C), public :: ngap_uns_opts_t
      real(c_double) :: strong_threshold,max_row_sum
      integer(c_int) :: pmax,max_coarse,max_levels,nu,kdim,graphs,verbose,precision
   end type ngap_uns_opts_t

   interface
      integer(c_int) function ngap_default_opts(o) bind(C,name='ngap_default_opts')
         import :: c_int,ngap_opts_t
         type(ngap_opts_t), intent(out) :: o
      end function ngap_default_opts
      integer(c_int) function ngap_create(h,g,nst,stc,method) bind(C,name='ngap_create')
         import :: c_int,c_ptr,ngap_grid_t
         type(c_ptr), intent(out) :: h
         type(ngap_grid_t), intent(in) :: g
         integer(c_int), value :: nst
         integer(c_int), intent(in) :: stc(*)
         integer(c_int), value :: method
      end function ngap_create
      integer(c_int) function ngap_init(h) bind(C,name='ngap_init')
         import :: c_int,c_ptr
         type(c_ptr), value :: h
      end function ngap_init
      integer(c_int) function ngap_setup(h,opr,maxit,rcvg,maxlevel,o) bind(C,name='ngap_setup')
         import :: c_int,c_ptr,c_double,ngap_opts_t
         type(c_ptr), value :: h
         real(c_double), intent(in) :: opr(*)
         integer(c_int), value :: maxit
         real(c_double), value :: rcvg
         integer(c_int), value :: maxlevel
         type(ngap_opts_t), intent(in) :: o
      end function ngap_setup
      integer(c_int) function ngap_solve(h,rhs,sol,it,rerr,aerr) bind(C,name='ngap_solve')
         import :: c_int,c_ptr,c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: rhs(*)
         real(c_double), intent(inout) :: sol(*)
         integer(c_int), intent(out) :: it
         real(c_double), intent(out) :: rerr,aerr
      end function ngap_solve
      integer(c_int) function ngap_set_tolerance(h,maxit,rcvg) bind(C,name='ngap_set_tolerance')
         import :: c_int,c_ptr,c_double
         type(c_ptr), value :: h
         integer(c_int), value :: maxit
         real(c_double), value :: rcvg
      end function ngap_set_tolerance
      !> Correction stage (learned guess / flexible-PCG stage): fn is c_funloc of a
      !> bind(C) function with the signature of ngap_stage_fn, e.g. ngapnet_stage
      integer(c_int) function ngap_set_stage(h,fn,ctx,flags,precond_iters) bind(C,name='ngap_set_stage')
         import :: c_int,c_ptr,c_funptr
         type(c_ptr), value :: h
         type(c_funptr), value :: fn
         type(c_ptr), value :: ctx
         integer(c_int), value :: flags,precond_iters
      end function ngap_set_stage
      integer(c_int) function ngap_destroy(h) bind(C,name='ngap_destroy')
         import :: c_int,c_ptr
         type(c_ptr), value :: h
      end function ngap_destroy
      integer(c_int) function ngap_last_stats(h,s) bind(C,name='ngap_last_stats')
         import :: c_int,c_ptr,ngap_stats_t
         type(c_ptr), value :: h
         type(ngap_stats_t), intent(out) :: s
      end function ngap_last_stats
      integer(c_int) function ngap_fft_create(f,g,nst,stc) bind(C,name='ngap_fft_create')
         import :: c_int,c_ptr,ngap_grid_t
         type(c_ptr), intent(out) :: f
         type(ngap_grid_t), intent(in) :: g
         integer(c_int), value :: nst
         integer(c_int), intent(in) :: stc(*)
      end function ngap_fft_create
      integer(c_int) function ngap_fft_setup(f,opr,circ_tol) bind(C,name='ngap_fft_setup')
         import :: c_int,c_ptr,c_double
         type(c_ptr), value :: f
         real(c_double), intent(in) :: opr(*)
         real(c_double), value :: circ_tol
      end function ngap_fft_setup
      integer(c_int) function ngap_fft_solve(f,rhs,sol) bind(C,name='ngap_fft_solve')
         import :: c_int,c_ptr,c_double
         type(c_ptr), value :: f
         real(c_double), intent(in) :: rhs(*)
         real(c_double), intent(inout) :: sol(*)
      end function ngap_fft_solve
      integer(c_int) function ngap_fft_destroy(f) bind(C,name='ngap_fft_destroy')
         import :: c_int,c_ptr
         type(c_ptr), value :: f
      end function ngap_fft_destroy
      integer(c_int) function ngap_uns_default_opts(o) bind(C,name='ngap_uns_default_opts')
         import :: c_int,ngap_uns_opts_t
         type(ngap_uns_opts_t), intent(out) :: o
      end function ngap_uns_default_opts
      integer(c_int) function ngap_uns_create(h,g,nst,stc,method) bind(C,name='ngap_uns_create')
         import :: c_int,c_ptr,ngap_grid_t
         type(c_ptr), intent(out) :: h
         type(ngap_grid_t), intent(in) :: g
         integer(c_int), value :: nst
         integer(c_int), intent(in) :: stc(*)
         integer(c_int), value :: method
      end function ngap_uns_create
      integer(c_int) function ngap_uns_init(h,opr) bind(C,name='ngap_uns_init')
         import :: c_int,c_ptr,c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: opr(*)
      end function ngap_uns_init
      integer(c_int) function ngap_uns_setup(h,opr,maxit,rcvg,o) bind(C,name='ngap_uns_setup')
         import :: c_int,c_ptr,c_double,ngap_uns_opts_t
         type(c_ptr), value :: h
         real(c_double), intent(in) :: opr(*)
         integer(c_int), value :: maxit
         real(c_double), value :: rcvg
         type(ngap_uns_opts_t), intent(in) :: o
      end function ngap_uns_setup
      integer(c_int) function ngap_uns_solve(h,rhs,sol,it,rerr,aerr) bind(C,name='ngap_uns_solve')
         import :: c_int,c_ptr,c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: rhs(*)
         real(c_double), intent(inout) :: sol(*)
         integer(c_int), intent(out) :: it
         real(c_double), intent(out) :: rerr,aerr
      end function ngap_uns_solve
      integer(c_int) function ngap_uns_destroy(h) bind(C,name='ngap_uns_destroy')
         import :: c_int,c_ptr
         type(c_ptr), value :: h
      end function ngap_uns_destroy
      integer(c_long_long) function ngap_uns_rows(h) bind(C,name='ngap_uns_rows')
         import :: c_long_long,c_ptr
         type(c_ptr), value :: h
      end function ngap_uns_rows
      integer(c_int) function ngap_uns_index(h,ind) bind(C,name='ngap_uns_index')
         import :: c_int,c_ptr
         type(c_ptr), value :: h
         integer(c_int), intent(out) :: ind(*)
      end function ngap_uns_index
      type(c_ptr) function ngap_strerror(code) bind(C,name='ngap_strerror')
         import :: c_int,c_ptr
         integer(c_int), value :: code
      end function ngap_strerror
   end interface

contains

   !> Message of a status code as a Fortran string
   function ngap_strerror_f(code) result(msg)
      integer, intent(in) :: code
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: s(:)
      integer :: n
      p=ngap_strerror(int(code,c_int))
      call c_f_pointer(p,s,[256])
      n=0
      do while (n.lt.256)
         if (s(n+1).eq.c_null_char) exit
         n=n+1
      end do
      allocate(character(len=n) :: msg)
      msg=transfer(s(1:n),msg)
   end function ngap_strerror_f

end module ngap_c
