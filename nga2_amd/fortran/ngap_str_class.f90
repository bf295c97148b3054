!> MI355X structured solver behind NGA2's linsol interface.
!> Drop-in for hypre_str with methods pcg_pfmg and gmres_pfmg (src/solver/hypre_str_class.f90
!> in NGA2): same constructor ngap_str(cfg,name,method,nst), same public
!> fields (opr, rhs, sol, stc, stmap, maxit, rcvg, it, rerr, aerr, maxlevel,
!> setup_done) and the same init/setup/solve/destroy/log/print/print_short
!> lifecycle. A driver switches by replacing hypre_str(...) with ngap_str(...)
!> (e.g. examples/channel/src/simulation.f90:137).
!> The numerics run in libngapsolve.so (C ABI, ngap_c); this module only
!> passes the arrays the caller owns.
module ngap_str_class
   use, intrinsic :: iso_c_binding
   use precision,    only: WP
   use config_class, only: config
   use string,       only: str_medium
   use linsol_class, only: linsol
   use ngap_c
   implicit none
   private

   public :: ngap_str
   public :: pfmg,pcg_pfmg,gmres_pfmg,smg,pcg_smg,gmres_smg,ngap_fp64,ngap_mixed
   public :: ngap_stage_guess,ngap_stage_precond

   !> ngap_str object definition
   type, extends(linsol) :: ngap_str
      integer :: maxlevel                                 !< Maximum number of multigrid levels (0: no cap)
      integer :: precision=ngap_fp64                      !< ngap_fp64, or ngap_mixed (FP32 V-cycle under FP64 PCG/GMRES)
      integer :: kdim=5                                   !< GMRES restart length (gmres_pfmg; HYPRE's default 5)
      type(c_ptr), private :: handle=c_null_ptr           !< libngapsolve handle (grid, stencil, device data)
      type(ngap_stats_t) :: stats                         !< statistics of the last solve
   contains
      procedure :: print_short=>ngap_str_print_short
      procedure :: print=>ngap_str_print
      procedure :: log=>ngap_str_log
      procedure :: init=>ngap_str_init
      procedure :: setup=>ngap_str_setup
      procedure :: solve=>ngap_str_solve
      procedure :: destroy=>ngap_str_destroy
      procedure :: set_stage=>ngap_str_set_stage
   end type ngap_str

   interface ngap_str
      procedure ngap_str_from_args
   end interface ngap_str

contains

   !> Constructor (mirrors hypre_str_from_args, hypre_str_class.f90:68-103)
   function ngap_str_from_args(cfg,name,method,nst) result(self)
      implicit none
      type(ngap_str) :: self
      class(config), target, intent(in) :: cfg
      character(len=*), intent(in) :: name
      integer, intent(in) :: method
      integer, intent(in) :: nst
      self%cfg=>cfg
      self%name=trim(adjustl(name))
      self%method=method
      self%nst=nst
      allocate(self%stc(1:self%nst,1:3)); self%stc=0
      allocate(self%opr(self%nst,self%cfg%imino_:self%cfg%imaxo_,self%cfg%jmino_:self%cfg%jmaxo_,self%cfg%kmino_:self%cfg%kmaxo_)); self%opr=0.0_WP
      allocate(self%rhs(         self%cfg%imino_:self%cfg%imaxo_,self%cfg%jmino_:self%cfg%jmaxo_,self%cfg%kmino_:self%cfg%kmaxo_)); self%rhs=0.0_WP
      allocate(self%sol(         self%cfg%imino_:self%cfg%imaxo_,self%cfg%jmino_:self%cfg%jmaxo_,self%cfg%kmino_:self%cfg%kmaxo_)); self%sol=0.0_WP
      self%it=0
      self%aerr=0.0_WP
      self%rerr=0.0_WP
      self%maxlevel=0
      self%setup_done=.false.
   end function ngap_str_from_args

   !> Abort with the library's message when a call fails
   subroutine check(rc,where)
      use messager, only: die
      implicit none
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc.ne.0) call die('[ngap_str '//where//'] '//ngap_strerror_f(int(rc)))
   end subroutine check

   !> Grid and stencil (mirrors hypre_str_init, :109-145)
   subroutine ngap_str_init(this)
      implicit none
      class(ngap_str), intent(inout) :: this
      integer :: st,stx1,stx2,sty1,sty2,stz1,stz2
      type(ngap_grid_t) :: g
      integer(c_int), dimension(:), allocatable :: stc
      ! inverse stencil map
      stx1=minval(this%stc(:,1)); stx2=maxval(this%stc(:,1))
      sty1=minval(this%stc(:,2)); sty2=maxval(this%stc(:,2))
      stz1=minval(this%stc(:,3)); stz2=maxval(this%stc(:,3))
      if (allocated(this%stmap)) deallocate(this%stmap)
      allocate(this%stmap(stx1:stx2,sty1:sty2,stz1:stz2)); this%stmap=0
      do st=1,this%nst
         this%stmap(this%stc(st,1),this%stc(st,2),this%stc(st,3))=st
      end do
      ! grid of this rank (one GPU holds the whole grid)
      g%nx=this%cfg%nx; g%ny=this%cfg%ny; g%nz=this%cfg%nz
      g%imin_=this%cfg%imin_; g%imax_=this%cfg%imax_
      g%jmin_=this%cfg%jmin_; g%jmax_=this%cfg%jmax_
      g%kmin_=this%cfg%kmin_; g%kmax_=this%cfg%kmax_
      g%imino_=this%cfg%imino_; g%imaxo_=this%cfg%imaxo_
      g%jmino_=this%cfg%jmino_; g%jmaxo_=this%cfg%jmaxo_
      g%kmino_=this%cfg%kmino_; g%kmaxo_=this%cfg%kmaxo_
      g%xper=merge(1,0,this%cfg%xper); g%yper=merge(1,0,this%cfg%yper); g%zper=merge(1,0,this%cfg%zper)
      allocate(stc(3*this%nst)); stc=reshape(this%stc,[3*this%nst])
      if (.not.c_associated(this%handle)) call check(ngap_create(this%handle,g,int(this%nst,c_int),stc,int(this%method,c_int)),'init')
      call check(ngap_init(this%handle),'init')
      deallocate(stc)
   end subroutine ngap_str_init

   !> Operator deep copy and hierarchy build (mirrors hypre_str_setup, :149-318)
   subroutine ngap_str_setup(this)
      implicit none
      class(ngap_str), intent(inout) :: this
      type(ngap_opts_t) :: o
      if (.not.c_associated(this%handle)) call this%init()
      call check(ngap_default_opts(o),'setup')
      o%precision=int(this%precision,c_int)
      o%kdim=int(this%kdim,c_int)
      call check(ngap_setup(this%handle,this%opr,int(this%maxit,c_int),real(this%rcvg,c_double),int(this%maxlevel,c_int),o),'setup')
      this%setup_done=.true.
   end subroutine ngap_str_setup

   !> Solve (mirrors hypre_str_solve, :322-386): sol holds the initial guess
   subroutine ngap_str_solve(this)
      use messager, only: die
      use param,    only: verbose
      implicit none
      class(ngap_str), intent(inout) :: this
      integer(c_int) :: it
      real(c_double) :: rerr,aerr
      if (.not.this%setup_done) call die('[ngap_str solve] Solver has not been setup.')
      this%it=-1; this%aerr=huge(1.0_WP); this%rerr=huge(1.0_WP)
      call check(ngap_solve(this%handle,this%rhs,this%sol,it,rerr,aerr),'solve')
      this%it=it; this%rerr=rerr; this%aerr=aerr
      call check(ngap_last_stats(this%handle,this%stats),'solve')
      call this%cfg%sync(this%sol)
      if (verbose.gt.0) call this%log
      if (verbose.gt.1) call this%print_short
   end subroutine ngap_str_solve

   !> Attach a correction stage (no counterpart in hypre_str): fn is c_funloc of
   !> a bind(C) function with the signature of ngap_stage_fn (e.g. ngapnet_stage
   !> of libngapnet.so) and ctx its context; flags ngap_stage_guess and/or
   !> ngap_stage_precond. fn=c_null_funptr removes the stage. Call after init.
   subroutine ngap_str_set_stage(this,fn,ctx,flags,precond_iters)
      use messager, only: die
      implicit none
      class(ngap_str), intent(inout) :: this
      type(c_funptr), intent(in) :: fn
      type(c_ptr), intent(in) :: ctx
      integer, intent(in) :: flags,precond_iters
      if (.not.c_associated(this%handle)) call die('[ngap_str set_stage] Solver has not been initialised.')
      call check(ngap_set_stage(this%handle,fn,ctx,int(flags,c_int),int(precond_iters,c_int)),'set_stage')
   end subroutine ngap_str_set_stage

   !> Release the device data (mirrors hypre_str_destroy, :390-422)
   subroutine ngap_str_destroy(this)
      implicit none
      class(ngap_str), intent(inout) :: this
      if (c_associated(this%handle)) call check(ngap_destroy(this%handle),'destroy')
      this%handle=c_null_ptr
      this%setup_done=.false.
   end subroutine ngap_str_destroy

   !> What log and print report: the last setup (levels, precision, method)
   !> and the last solve (iterations, the true residual with the exact
   !> operator beside the recursive one, the C-norm ratio HYPRE's PCG stops
   !> on, fallbacks, times)
   subroutine ngap_str_report(this,line,n)
      implicit none
      class(ngap_str), intent(in) :: this
      character(len=*), dimension(:), intent(out) :: line
      integer, intent(out) :: n
      character(len=12) :: meth,prec,state
      meth='pcg_pfmg'
      if (this%method.eq.gmres_pfmg) meth='gmres_pfmg'
      prec='fp64'
      if (this%stats%precision.eq.ngap_mixed) prec='mixed'
      state='converged'
      if (this%stats%status.eq.ngap_status_maxit) state='at maxit'
      if (this%stats%status.eq.ngap_status_stagnated) state='stagnated'
      line=''
      write(line(1),'("ngap_str ",a," on ",a,": ",a," (",a,"), ",i0," multigrid levels")') &
      &     trim(this%name),trim(this%cfg%name),trim(meth),trim(prec),this%stats%nlevels
      write(line(2),'("   ",i0," iterations of ",i0,", ",a,"; rcvg ",es9.2)') this%it,this%maxit,trim(state),this%rcvg
      write(line(3),'("   |b-Ax|/|b| exact operator ",es10.3,", recursive ",es10.3,", C-norm ratio ",es10.3)') &
      &     this%stats%true_rerr,this%stats%rerr,this%stats%cnorm_ratio
      write(line(4),'("   |b-Ax| ",es10.3,"; fallback flags ",i0,", residual replacements ",i0)') &
      &     this%aerr,this%stats%fallback,this%stats%restarts
      write(line(5),'("   setup ",f0.3," ms, solve ",f0.3," ms")') this%stats%setup_ms,this%stats%solve_ms
      n=5
   end subroutine ngap_str_report

   subroutine ngap_str_log(this)
      use string,   only: str_long
      use messager, only: log
      implicit none
      class(ngap_str), intent(in) :: this
      character(len=str_long), dimension(5) :: line
      integer :: n,l
      if (.not.this%cfg%amRoot) return
      call ngap_str_report(this,line,n)
      do l=1,n
         call log(line(l))
      end do
   end subroutine ngap_str_log

   subroutine ngap_str_print(this)
      use, intrinsic :: iso_fortran_env, only: output_unit
      use string, only: str_long
      implicit none
      class(ngap_str), intent(in) :: this
      character(len=str_long), dimension(5) :: line
      integer :: n,l
      if (.not.this%cfg%amRoot) return
      call ngap_str_report(this,line,n)
      do l=1,n
         write(output_unit,'(a)') trim(line(l))
      end do
   end subroutine ngap_str_print

   subroutine ngap_str_print_short(this)
      use, intrinsic :: iso_fortran_env, only: output_unit
      implicit none
      class(ngap_str), intent(in) :: this
      if (.not.this%cfg%amRoot) return
      write(output_unit,'("ngap_str ",a,": it ",i0,"/",i0,", |b-Ax|/|b| ",es10.3," (rcvg ",es9.2,"), ",i0," levels")') &
      &     trim(this%name),this%it,this%maxit,this%stats%true_rerr,this%rcvg,this%stats%nlevels
   end subroutine ngap_str_print_short

end module ngap_str_class
