!> Test scaffolding: a one-rank structured config with the fields and the
!> sync the ngap_str shim reads (NGA2: src/config/config_class.f90 on top of
!> src/grid/pgrid_class.f90). Uniform mesh, ghost width no, periodic wrap in
!> periodic directions.
module config_class
   use precision, only: WP
   use string,    only: str_medium
   implicit none
   private
   public :: config

   type :: config
      character(len=str_medium) :: name='config'
      integer :: nx,ny,nz,no
      integer :: imin_,imax_,jmin_,jmax_,kmin_,kmax_
      integer :: imino_,imaxo_,jmino_,jmaxo_,kmino_,kmaxo_
      logical :: xper,yper,zper
      logical :: amRoot=.true.
      real(WP) :: dx,dy,dz
   contains
      procedure :: sync=>config_sync
   end type config

   interface config
      procedure config_from_args
   end interface config

contains

   function config_from_args(nx,ny,nz,no,Lx,Ly,Lz,xper,yper,zper) result(self)
      type(config) :: self
      integer, intent(in) :: nx,ny,nz,no
      real(WP), intent(in) :: Lx,Ly,Lz
      logical, intent(in) :: xper,yper,zper
      self%nx=nx; self%ny=ny; self%nz=nz; self%no=no
      self%imin_=1; self%imax_=nx; self%jmin_=1; self%jmax_=ny; self%kmin_=1; self%kmax_=nz
      self%imino_=1-no; self%imaxo_=nx+no; self%jmino_=1-no; self%jmaxo_=ny+no; self%kmino_=1-no; self%kmaxo_=nz+no
      self%xper=xper; self%yper=yper; self%zper=zper
      self%dx=Lx/real(nx,WP); self%dy=Ly/real(ny,WP); self%dz=Lz/real(nz,WP)
   end function config_from_args

   !> Ghost fill: periodic copy in periodic directions, nothing otherwise
   subroutine config_sync(this,A)
      class(config), intent(in) :: this
      real(WP), dimension(this%imino_:,this%jmino_:,this%kmino_:), intent(inout) :: A
      integer :: g
      do g=1,this%no
         if (this%xper) then
            A(this%imin_-g,:,:)=A(this%imax_-g+1,:,:)
            A(this%imax_+g,:,:)=A(this%imin_+g-1,:,:)
         end if
         if (this%yper) then
            A(:,this%jmin_-g,:)=A(:,this%jmax_-g+1,:)
            A(:,this%jmax_+g,:)=A(:,this%jmin_+g-1,:)
         end if
         if (this%zper) then
            A(:,:,this%kmin_-g)=A(:,:,this%kmax_-g+1)
            A(:,:,this%kmax_+g)=A(:,:,this%kmin_+g-1)
         end if
      end do
   end subroutine config_sync

end module config_class
