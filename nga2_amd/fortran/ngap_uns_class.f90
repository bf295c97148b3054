!> Unstructured (IJ) linear solver on the GPU behind NGA2's linsol interface.
!> Drop-in for hypre_uns (src/solver/hypre_uns_class.f90 in NGA2): same
!> constructor ngap_uns(cfg,name,method,nst), the same method constants
!> (amg, pcg_amg, gmres_amg, bicgstab_amg, pcg, pcg_parasail, gmres,
!> gmres_pilut), the same public fields (maxlevel, ind, ncell, ncell_) and the
!> same init/setup/solve/destroy/log/print/print_short lifecycle. init numbers
!> the cells with |opr(1)| > 10 eps (prep_umap), so opr(1) must be set before
!> it, as for the reference. pcg_parasail and gmres_pilut are not provided:
!> setup stops with a message for them. A driver switches by replacing
!> hypre_uns(...) with ngap_uns(...).
!> The matrix, the multigrid hierarchy and every iteration live in
!> libngapuns.so (C ABI, ngap_c); this module only passes the arrays the
!> caller owns.
module ngap_uns_class
   use, intrinsic :: iso_c_binding
   use precision,    only: WP
   use config_class, only: config
   use linsol_class, only: linsol
   use ngap_c,       only: ngap_grid_t,ngap_uns_opts_t,ngap_uns_default_opts,ngap_uns_create,ngap_uns_init, &
   &                       ngap_uns_setup,ngap_uns_solve,ngap_uns_destroy,ngap_uns_rows,ngap_uns_index,ngap_strerror_f
   implicit none
   private

   public :: ngap_uns

   ! hypre_uns_class.f90:17-24
   integer, parameter, public ::          amg=1
   integer, parameter, public ::      pcg_amg=2
   integer, parameter, public ::    gmres_amg=3
   integer, parameter, public :: bicgstab_amg=4
   integer, parameter, public :: pcg         =5
   integer, parameter, public :: pcg_parasail=6
   integer, parameter, public ::  gmres      =7
   integer, parameter, public ::  gmres_pilut=8

   type, extends(linsol) :: ngap_uns
      integer :: maxlevel                                !< kept for the reference's interface (unused there too)
      integer,  dimension(:,:,:), allocatable :: ind     !< Unique index of each participating cell, -1 elsewhere
      integer :: ind_min,ind_max                         !< Local min and max indices
      integer :: ncell_,ncell                            !< Local and global participating cells
      type(ngap_uns_opts_t) :: opts                      !< build-specific knobs, hypre_uns's values by default
      type(c_ptr), private :: handle=c_null_ptr          !< libngapuns handle
   contains
      procedure :: print_short=>ngap_uns_print_short
      procedure :: print=>ngap_uns_print
      procedure :: log=>ngap_uns_log
      procedure :: init=>ngap_uns_init_
      procedure :: setup=>ngap_uns_setup_
      procedure :: solve=>ngap_uns_solve_
      procedure :: destroy=>ngap_uns_destroy_
   end type ngap_uns

   interface ngap_uns
      procedure ngap_uns_from_args
   end interface ngap_uns

contains

   !> Constructor (mirrors hypre_uns_from_args, hypre_uns_class.f90:84-119)
   function ngap_uns_from_args(cfg,name,method,nst) result(self)
      implicit none
      type(ngap_uns) :: self
      class(config), target, intent(in) :: cfg
      character(len=*), intent(in) :: name
      integer, intent(in) :: method
      integer, intent(in) :: nst
      integer(c_int) :: rc
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
      self%ncell=0; self%ncell_=0; self%ind_min=0; self%ind_max=-1
      rc=ngap_uns_default_opts(self%opts)
      self%setup_done=.false.
   end function ngap_uns_from_args

   subroutine check(rc,where)
      use messager, only: die
      implicit none
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc.ne.0) call die('[ngap_uns '//where//'] '//ngap_strerror_f(int(rc)))
   end subroutine check

   !> Stencil map and cell numbering (mirrors hypre_uns_init, :124-149, and prep_umap, :562-607)
   subroutine ngap_uns_init_(this)
      implicit none
      class(ngap_uns), intent(inout) :: this
      integer :: st,stx1,stx2,sty1,sty2,stz1,stz2
      type(ngap_grid_t) :: g
      integer(c_int), dimension(:), allocatable :: stc
      integer(c_int), dimension(:,:,:), allocatable :: cind
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
      if (c_associated(this%handle)) call check(ngap_uns_destroy(this%handle),'init')
      this%handle=c_null_ptr
      call check(ngap_uns_create(this%handle,g,int(this%nst,c_int),stc,int(this%method,c_int)),'init')
      deallocate(stc)
      call check(ngap_uns_init(this%handle,this%opr),'init')
      if (.not.allocated(this%ind)) allocate(this%ind(this%cfg%imino_:this%cfg%imaxo_,this%cfg%jmino_:this%cfg%jmaxo_,this%cfg%kmino_:this%cfg%kmaxo_))
      allocate(cind(this%cfg%imino_:this%cfg%imaxo_,this%cfg%jmino_:this%cfg%jmaxo_,this%cfg%kmino_:this%cfg%kmaxo_))
      call check(ngap_uns_index(this%handle,cind),'init')
      this%ind=int(cind)
      deallocate(cind)
      this%ncell_=int(ngap_uns_rows(this%handle))
      this%ncell=this%ncell_
      this%ind_min=0
      this%ind_max=this%ncell_-1
   end subroutine ngap_uns_init_

   !> Matrix and hierarchy (mirrors hypre_uns_setup, :153-426)
   subroutine ngap_uns_setup_(this)
      use messager, only: die
      implicit none
      class(ngap_uns), intent(inout) :: this
      if (.not.c_associated(this%handle)) call die('[ngap_uns setup] init has not been called')
      if (this%method.eq.pcg_parasail.or.this%method.eq.gmres_pilut) call die('[ngap_uns setup] method not provided on the GPU (ParaSails, PILUT)')
      this%setup_done=.false.
      call check(ngap_uns_setup(this%handle,this%opr,int(this%maxit,c_int),real(this%rcvg,c_double),this%opts),'setup')
      this%setup_done=.true.
   end subroutine ngap_uns_setup_

   !> Solve (mirrors hypre_uns_solve, :430-500)
   subroutine ngap_uns_solve_(this)
      use messager, only: die
      use param,    only: verbose
      implicit none
      class(ngap_uns), intent(inout) :: this
      integer(c_int) :: it
      real(c_double) :: rerr,aerr
      if (.not.this%setup_done) call die('[ngap_uns solve] Solver has not been setup.')
      this%it=-1; this%aerr=huge(1.0_WP); this%rerr=huge(1.0_WP)
      call check(ngap_uns_solve(this%handle,this%rhs,this%sol,it,rerr,aerr),'solve')
      this%it=int(it); this%rerr=real(rerr,WP); this%aerr=real(aerr,WP)
      call this%cfg%sync(this%sol)
      if (verbose.gt.0) call this%log
      if (verbose.gt.1) call this%print_short
   end subroutine ngap_uns_solve_

   !> Release the matrix, hierarchy and vectors (mirrors hypre_uns_destroy, :504-544)
   subroutine ngap_uns_destroy_(this)
      implicit none
      class(ngap_uns), intent(inout) :: this
      if (c_associated(this%handle)) call check(ngap_uns_destroy(this%handle),'destroy')
      this%handle=c_null_ptr
      this%setup_done=.false.
   end subroutine ngap_uns_destroy_

   !> What log and print report: the method, the participating rows and the
   !> last solve's iterations and residuals (2-norm stopping rule)
   subroutine ngap_uns_report(this,line,n)
      implicit none
      class(ngap_uns), intent(in) :: this
      character(len=*), dimension(:), intent(out) :: line
      integer, intent(out) :: n
      character(len=14), dimension(8), parameter :: mname=[character(len=14) :: 'amg','pcg_amg','gmres_amg', &
      &     'bicgstab_amg','pcg','pcg_parasail','gmres','gmres_pilut']
      character(len=14) :: meth
      integer(c_long_long) :: rows
      meth='unknown'
      if (this%method.ge.1.and.this%method.le.8) meth=mname(this%method)
      rows=0
      if (c_associated(this%handle)) rows=ngap_uns_rows(this%handle)
      line=''
      write(line(1),'("ngap_uns ",a," on ",a,": ",a,", ",i0," participating cells")') &
      &     trim(this%name),trim(this%cfg%name),trim(meth),rows
      write(line(2),'("   ",i0," iterations of ",i0,"; |b-Ax|/|b| ",es10.3," (rcvg ",es9.2,"), |b-Ax| ",es10.3)') &
      &     this%it,this%maxit,this%rerr,this%rcvg,this%aerr
      n=2
   end subroutine ngap_uns_report

   subroutine ngap_uns_log(this)
      use string,   only: str_long
      use messager, only: log
      implicit none
      class(ngap_uns), intent(in) :: this
      character(len=str_long), dimension(2) :: line
      integer :: n,l
      if (.not.this%cfg%amRoot) return
      call ngap_uns_report(this,line,n)
      do l=1,n
         call log(line(l))
      end do
   end subroutine ngap_uns_log

   subroutine ngap_uns_print(this)
      use, intrinsic :: iso_fortran_env, only: output_unit
      use string, only: str_long
      implicit none
      class(ngap_uns), intent(in) :: this
      character(len=str_long), dimension(2) :: line
      integer :: n,l
      if (.not.this%cfg%amRoot) return
      call ngap_uns_report(this,line,n)
      do l=1,n
         write(output_unit,'(a)') trim(line(l))
      end do
   end subroutine ngap_uns_print

   subroutine ngap_uns_print_short(this)
      use, intrinsic :: iso_fortran_env, only: output_unit
      implicit none
      class(ngap_uns), intent(in) :: this
      if (.not.this%cfg%amRoot) return
      write(output_unit,'("ngap_uns ",a,": it ",i0,"/",i0,", |b-Ax|/|b| ",es10.3," (rcvg ",es9.2,")")') &
      &     trim(this%name),this%it,this%maxit,this%rerr,this%rcvg
   end subroutine ngap_uns_print_short

end module ngap_uns_class
