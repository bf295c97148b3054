!> FFT pressure solver on the GPU behind NGA2's linsol interface.
!> Drop-in for fourier3d (src/solver/fourier3d_class.f90 in NGA2): same
!> constructor ngap_fourier3d(cfg,name,nst) with the same checks (periodic and
!> uniform in every direction with more than one cell, at least one direction
!> not decomposed), the same public fields and the same
!> init/setup/solve/destroy/log/print/print_short lifecycle; it=1 and
!> rerr=aerr=0 as the reference leaves them. A driver switches by replacing
!> fourier3d(...) with ngap_fourier3d(...).
!> The transforms and kernels run in libngapfft.so (C ABI, ngap_c); this
!> module only passes the arrays the caller owns.
module ngap_fourier3d_class
   use, intrinsic :: iso_c_binding
   use precision,    only: WP
   use config_class, only: config
   use linsol_class, only: linsol
   use ngap_c,       only: ngap_grid_t,ngap_fft_create,ngap_fft_setup,ngap_fft_solve,ngap_fft_destroy,ngap_strerror_f
   implicit none
   private

   public :: ngap_fourier3d

   type, extends(linsol) :: ngap_fourier3d
      type(c_ptr), private :: handle=c_null_ptr           !< libngapfft handle (plans, 1/F on the device)
   contains
      procedure :: print_short=>ngap_fourier3d_print_short
      procedure :: print=>ngap_fourier3d_print
      procedure :: log=>ngap_fourier3d_log
      procedure :: init=>ngap_fourier3d_init
      procedure :: setup=>ngap_fourier3d_setup
      procedure :: solve=>ngap_fourier3d_solve
      procedure :: destroy=>ngap_fourier3d_destroy
   end type ngap_fourier3d

   interface ngap_fourier3d
      procedure ngap_fourier3d_from_args
   end interface ngap_fourier3d

contains

   !> Constructor (mirrors fourier3d_from_args, fourier3d_class.f90:109-165)
   function ngap_fourier3d_from_args(cfg,name,nst) result(self)
      use messager, only: die
      implicit none
      type(ngap_fourier3d) :: self
      class(config), target, intent(in) :: cfg
      character(len=*), intent(in) :: name
      integer, intent(in) :: nst
      integer :: ndim,ndcp
      self%cfg=>cfg
      self%name=trim(adjustl(name))
      self%method=0
      self%nst=nst
      allocate(self%stc(1:self%nst,1:3)); self%stc=0
      allocate(self%opr(self%nst,self%cfg%imino_:self%cfg%imaxo_,self%cfg%jmino_:self%cfg%jmaxo_,self%cfg%kmino_:self%cfg%kmaxo_)); self%opr=0.0_WP
      allocate(self%rhs(         self%cfg%imino_:self%cfg%imaxo_,self%cfg%jmino_:self%cfg%jmaxo_,self%cfg%kmino_:self%cfg%kmaxo_)); self%rhs=0.0_WP
      allocate(self%sol(         self%cfg%imino_:self%cfg%imaxo_,self%cfg%jmino_:self%cfg%jmaxo_,self%cfg%kmino_:self%cfg%kmaxo_)); self%sol=0.0_WP
      self%it=1
      self%aerr=0.0_WP
      self%rerr=0.0_WP
      self%setup_done=.false.
      if (self%cfg%nx.gt.1.and..not.(self%cfg%xper.and.self%cfg%uniform_x)) call die('[ngap_fourier3d constructor] x-direction needs to be periodic and uniform')
      if (self%cfg%ny.gt.1.and..not.(self%cfg%yper.and.self%cfg%uniform_y)) call die('[ngap_fourier3d constructor] y-direction needs to be periodic and uniform')
      if (self%cfg%nz.gt.1.and..not.(self%cfg%zper.and.self%cfg%uniform_z)) call die('[ngap_fourier3d constructor] z-direction needs to be periodic and uniform')
      ndim=count([self%cfg%nx,self%cfg%ny,self%cfg%nz].gt.1)
      ndcp=count([self%cfg%npx,self%cfg%npy,self%cfg%npz].gt.1)
      if (ndcp.ge.ndim) call die('[ngap_fourier3d constructor] Need at least one NON-decomposed direction')
   end function ngap_fourier3d_from_args

   subroutine check(rc,where)
      use messager, only: die
      implicit none
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc.ne.0) call die('[ngap_fourier3d '//where//'] '//ngap_strerror_f(int(rc)))
   end subroutine check

   !> Stencil map, grid and transform plans (mirrors fourier3d_init, :171-216)
   subroutine ngap_fourier3d_init(this)
      implicit none
      class(ngap_fourier3d), intent(inout) :: this
      integer :: st,stx1,stx2,sty1,sty2,stz1,stz2
      type(ngap_grid_t) :: g
      integer(c_int), dimension(:), allocatable :: stc
      stx1=minval(this%stc(:,1)); stx2=maxval(this%stc(:,1))
      sty1=minval(this%stc(:,2)); sty2=maxval(this%stc(:,2))
      stz1=minval(this%stc(:,3)); stz2=maxval(this%stc(:,3))
      if (allocated(this%stmap)) deallocate(this%stmap)
      allocate(this%stmap(stx1:stx2,sty1:sty2,stz1:stz2)); this%stmap=0
      do st=1,this%nst
         this%stmap(this%stc(st,1),this%stc(st,2),this%stc(st,3))=st
      end do
      g%nx=this%cfg%nx; g%ny=this%cfg%ny; g%nz=this%cfg%nz
      g%imin_=this%cfg%imin_; g%imax_=this%cfg%imax_
      g%jmin_=this%cfg%jmin_; g%jmax_=this%cfg%jmax_
      g%kmin_=this%cfg%kmin_; g%kmax_=this%cfg%kmax_
      g%imino_=this%cfg%imino_; g%imaxo_=this%cfg%imaxo_
      g%jmino_=this%cfg%jmino_; g%jmaxo_=this%cfg%jmaxo_
      g%kmino_=this%cfg%kmino_; g%kmaxo_=this%cfg%kmaxo_
      g%xper=merge(1,0,this%cfg%xper); g%yper=merge(1,0,this%cfg%yper); g%zper=merge(1,0,this%cfg%zper)
      allocate(stc(3*this%nst)); stc=reshape(this%stc,[3*this%nst])
      if (.not.c_associated(this%handle)) call check(ngap_fft_create(this%handle,g,int(this%nst,c_int),stc),'init')
      deallocate(stc)
   end subroutine ngap_fourier3d_init

   !> Circulance check and 1/F (mirrors fourier3d_setup, :1103-1166)
   subroutine ngap_fourier3d_setup(this)
      implicit none
      class(ngap_fourier3d), intent(inout) :: this
      if (.not.c_associated(this%handle)) call this%init()
      this%setup_done=.false.
      call check(ngap_fft_setup(this%handle,this%opr,real(6.0_WP*epsilon(1.0_WP)/this%cfg%min_meshsize**4,c_double)),'setup')
      this%setup_done=.true.
   end subroutine ngap_fourier3d_setup

   !> Solve (mirrors fourier3d_solve, :1170-1202); the ghosts are filled by the library
   subroutine ngap_fourier3d_solve(this)
      use messager, only: die
      use param,    only: verbose
      implicit none
      class(ngap_fourier3d), intent(inout) :: this
      if (.not.this%setup_done) call die('[ngap_fourier3d solve] Solver has not been setup.')
      call check(ngap_fft_solve(this%handle,this%rhs,this%sol),'solve')
      if (verbose.gt.0) call this%log
      if (verbose.gt.1) call this%print_short
   end subroutine ngap_fourier3d_solve

   !> Release plans and device data (mirrors fourier3d_destroy, :1206-1221)
   subroutine ngap_fourier3d_destroy(this)
      implicit none
      class(ngap_fourier3d), intent(inout) :: this
      if (c_associated(this%handle)) call check(ngap_fft_destroy(this%handle),'destroy')
      this%handle=c_null_ptr
      this%setup_done=.false.
   end subroutine ngap_fourier3d_destroy

   !> The one line log and print report: a direct solve has no iterations
   !> or residual history, so the grid it transforms is what it states
   subroutine ngap_fourier3d_report(this,line)
      implicit none
      class(ngap_fourier3d), intent(in) :: this
      character(len=*), intent(out) :: line
      line=''
      write(line,'("ngap_fourier3d ",a," on ",a,": FFT solve of a circulant stencil on ",i0," x ",i0," x ",i0, &
      &     " cells, set up ",l1)') trim(this%name),trim(this%cfg%name),this%cfg%nx,this%cfg%ny,this%cfg%nz, &
      &     this%setup_done
   end subroutine ngap_fourier3d_report

   subroutine ngap_fourier3d_log(this)
      use string,   only: str_long
      use messager, only: log
      implicit none
      class(ngap_fourier3d), intent(in) :: this
      character(len=str_long) :: line
      if (.not.this%cfg%amRoot) return
      call ngap_fourier3d_report(this,line)
      call log(line)
   end subroutine ngap_fourier3d_log

   subroutine ngap_fourier3d_print(this)
      use, intrinsic :: iso_fortran_env, only: output_unit
      use string, only: str_long
      implicit none
      class(ngap_fourier3d), intent(in) :: this
      character(len=str_long) :: line
      if (.not.this%cfg%amRoot) return
      call ngap_fourier3d_report(this,line)
      write(output_unit,'(a)') trim(line)
   end subroutine ngap_fourier3d_print

   subroutine ngap_fourier3d_print_short(this)
      use, intrinsic :: iso_fortran_env, only: output_unit
      implicit none
      class(ngap_fourier3d), intent(in) :: this
      if (.not.this%cfg%amRoot) return
      write(output_unit,'("ngap_fourier3d ",a,": direct solve, ",i0," x ",i0," x ",i0)') trim(this%name), &
      &     this%cfg%nx,this%cfg%ny,this%cfg%nz
   end subroutine ngap_fourier3d_print_short

end module ngap_fourier3d_class
