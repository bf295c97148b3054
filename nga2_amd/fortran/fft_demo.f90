!> Self-checking Fortran driver of the ngap_fourier3d drop-in, written the way
!> an NGA2 driver uses fourier3d (examples/hit): construct, set stc, init, fill
!> opr, setup, fill rhs, solve. A manufactured periodic solution p* gives
!> rhs = A p*; the direct solve must return p* (mean-free) to 1e-10 with its
!> ghosts synced, on an anisotropic 3-D box and on a one-cell-thick 2-D box.
!> Exit status 0 when every case passes.
program ngap_fft_demo
   use precision,            only: WP
   use config_class,         only: config
   use ngap_fourier3d_class, only: ngap_fourier3d
   implicit none
   integer :: n,nfail
   character(len=32) :: arg
   n=32
   if (command_argument_count().ge.1) then
      call get_command_argument(1,arg); read(arg,*) n
   end if
   nfail=0
   call run_case('box 3d',n,n/2,n,1.0_WP,0.5_WP,2.0_WP,.true.,nfail)
   call run_case('box 2d',n,n,1,1.0_WP,1.0_WP,0.1_WP,.false.,nfail)
   if (nfail.gt.0) then
      write(*,'("NGAP_FFT_DEMO FAIL ",i0)') nfail
      error stop 1
   end if
   write(*,'("NGAP_FFT_DEMO OK")')

contains

   subroutine run_case(label,nx,ny,nz,Lx,Ly,Lz,zper,nfail)
      character(len=*), intent(in) :: label
      integer, intent(in) :: nx,ny,nz
      real(WP), intent(in) :: Lx,Ly,Lz
      logical, intent(in) :: zper
      integer, intent(inout) :: nfail
      type(config), target :: cfg
      type(ngap_fourier3d) :: fs
      real(WP), dimension(:,:,:), allocatable :: pstar,Ap
      real(WP) :: vol,cx,cy,cz,pi,x,y,z,err,nrm,m1,m2
      integer :: i,j,k,st
      pi=4.0_WP*atan(1.0_WP)
      cfg=config(nx,ny,nz,1,Lx,Ly,Lz,.true.,.true.,zper)
      cfg%name='fftdemo'
      fs=ngap_fourier3d(cfg=cfg,name='Pressure',nst=7)
      fs%stc(1,:)=[ 0, 0, 0]
      fs%stc(2,:)=[+1, 0, 0]
      fs%stc(3,:)=[-1, 0, 0]
      fs%stc(4,:)=[ 0,+1, 0]
      fs%stc(5,:)=[ 0,-1, 0]
      fs%stc(6,:)=[ 0, 0,+1]
      fs%stc(7,:)=[ 0, 0,-1]
      call fs%init()
      vol=cfg%dx*cfg%dy*cfg%dz
      cx=vol/cfg%dx**2; cy=vol/cfg%dy**2; cz=vol/cfg%dz**2
      if (nz.eq.1) cz=0.0_WP
      do k=cfg%kmin_,cfg%kmax_
         do j=cfg%jmin_,cfg%jmax_
            do i=cfg%imin_,cfg%imax_
               fs%opr(2:3,i,j,k)=-cx
               fs%opr(4:5,i,j,k)=-cy
               fs%opr(6:7,i,j,k)=-cz
               fs%opr(1,i,j,k)=-sum(fs%opr(2:7,i,j,k))
            end do
         end do
      end do
      call fs%setup()
      allocate(pstar(cfg%imino_:cfg%imaxo_,cfg%jmino_:cfg%jmaxo_,cfg%kmino_:cfg%kmaxo_)); pstar=0.0_WP
      allocate(Ap,mold=pstar); Ap=0.0_WP
      do k=cfg%kmin_,cfg%kmax_
         z=(real(k,WP)-0.5_WP)*cfg%dz
         do j=cfg%jmin_,cfg%jmax_
            y=(real(j,WP)-0.5_WP)*cfg%dy
            do i=cfg%imin_,cfg%imax_
               x=(real(i,WP)-0.5_WP)*cfg%dx
               pstar(i,j,k)=sin(2.0_WP*pi*x/Lx)*cos(4.0_WP*pi*y/Ly)+0.5_WP*cos(2.0_WP*pi*z/Lz)+0.1_WP*sin(6.0_WP*pi*x/Lx)
            end do
         end do
      end do
      call cfg%sync(pstar)
      do k=cfg%kmin_,cfg%kmax_
         do j=cfg%jmin_,cfg%jmax_
            do i=cfg%imin_,cfg%imax_
               Ap(i,j,k)=0.0_WP
               do st=1,7
                  Ap(i,j,k)=Ap(i,j,k)+fs%opr(st,i,j,k)*pstar(i+fs%stc(st,1),j+fs%stc(st,2),k+fs%stc(st,3))
               end do
            end do
         end do
      end do
      fs%rhs=Ap
      fs%sol=0.0_WP
      call fs%solve()
      m1=sum(fs%sol(cfg%imin_:cfg%imax_,cfg%jmin_:cfg%jmax_,cfg%kmin_:cfg%kmax_))/real(nx*ny*nz,WP)
      m2=sum(pstar(cfg%imin_:cfg%imax_,cfg%jmin_:cfg%jmax_,cfg%kmin_:cfg%kmax_))/real(nx*ny*nz,WP)
      err=sqrt(sum((fs%sol(cfg%imin_:cfg%imax_,cfg%jmin_:cfg%jmax_,cfg%kmin_:cfg%kmax_)-m1 &
      &            -pstar(cfg%imin_:cfg%imax_,cfg%jmin_:cfg%jmax_,cfg%kmin_:cfg%kmax_)+m2)**2))
      nrm=sqrt(sum((pstar(cfg%imin_:cfg%imax_,cfg%jmin_:cfg%jmax_,cfg%kmin_:cfg%kmax_)-m2)**2))
      err=err/nrm
      call fs%print_short()
      write(*,'(a," it=",i0," err=",es10.3," mean=",es10.3)') label,fs%it,err,m1
      if (err.gt.1.0e-10_WP.or.abs(m1).gt.1.0e-12_WP.or.fs%it.ne.1) nfail=nfail+1
      if (abs(fs%sol(cfg%imin_-1,1,1)-fs%sol(cfg%imax_,1,1)).gt.0.0_WP) nfail=nfail+1
      call fs%destroy()
      deallocate(pstar,Ap)
   end subroutine run_case

end program ngap_fft_demo
