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
      logical :: uniform_x=.true.,uniform_y=.true.,uniform_z=.true.  !< sgrid_class.f90:266-268 (uniform mesh here)
      integer :: npx=1,npy=1,npz=1                                    !< one rank
      real(WP) :: min_meshsize                                        !< config_class.f90:112-125
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
      ! geometric mean of the spacings of the directions with more than one cell
      block
         integer :: px,py,pz
         px=merge(0,1,nx.eq.1); py=merge(0,1,ny.eq.1); pz=merge(0,1,nz.eq.1)
         if (px+py+pz.gt.0) then
            self%min_meshsize=(self%dx**px*self%dy**py*self%dz**pz)**(1.0_WP/real(px+py+pz,WP))
         else
            self%min_meshsize=1.0_WP
         end if
      end block
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
