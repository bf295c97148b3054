!> Self-checking Fortran driver of the ngap_uns drop-in, written the way an
!> NGA2 driver uses hypre_uns: construct, set stc and opr (opr(1)=0 marks the
!> cells outside the system), init, setup, fill rhs and the guess, solve. A
!> periodic box holds a solid block whose cells are outside the system and
!> whose faces are homogeneous Neumann for the fluid. A manufactured fluid
!> solution p* gives rhs = A p*; each method must return p* (mean-free) to
!> 1e-8, leave the solid cells' sol untouched and sync the ghosts.
!> Exit status 0 when every case passes.
program ngap_uns_demo
   use precision,      only: WP
   use config_class,   only: config
   use ngap_uns_class, only: ngap_uns,amg,pcg_amg,gmres_amg,bicgstab_amg,pcg,gmres
   implicit none
   integer :: n,nfail,m
   integer, dimension(6) :: methods=[pcg_amg,gmres_amg,bicgstab_amg,amg,pcg,gmres]
   character(len=32) :: arg
   n=24
   if (command_argument_count().ge.1) then
      call get_command_argument(1,arg); read(arg,*) n
   end if
   nfail=0
   do m=1,size(methods)
      call run_case(methods(m),n,nfail)
   end do
   if (nfail.gt.0) then
      write(*,'("NGAP_UNS_DEMO FAIL ",i0)') nfail
      error stop 1
   end if
   write(*,'("NGAP_UNS_DEMO OK")')

contains

   logical function solid(i,j,k,n)
      integer, intent(in) :: i,j,k,n
      solid=(i.gt.n/4.and.i.le.n/2.and.j.gt.n/3.and.j.le.2*n/3.and.k.gt.n/4.and.k.le.3*n/4)
   end function solid

   subroutine run_case(method,n,nfail)
      integer, intent(in) :: method,n
      integer, intent(inout) :: nfail
      type(config), target :: cfg
      type(ngap_uns) :: us
      real(WP), dimension(:,:,:), allocatable :: pstar,Ap
      real(WP) :: h,pi,x,y,z,err,nrm,m1,m2,c
      integer :: i,j,k,st,nf
      integer, dimension(3,6) :: off
      pi=4.0_WP*atan(1.0_WP)
      cfg=config(n,n,n,1,1.0_WP,1.0_WP,1.0_WP,.true.,.true.,.true.)
      cfg%name='unsdemo'
      us=ngap_uns(cfg=cfg,name='Pressure',method=method,nst=7)
      us%stc(1,:)=[ 0, 0, 0]
      us%stc(2,:)=[+1, 0, 0]
      us%stc(3,:)=[-1, 0, 0]
      us%stc(4,:)=[ 0,+1, 0]
      us%stc(5,:)=[ 0,-1, 0]
      us%stc(6,:)=[ 0, 0,+1]
      us%stc(7,:)=[ 0, 0,-1]
      off=reshape([1,0,0, -1,0,0, 0,1,0, 0,-1,0, 0,0,1, 0,0,-1],[3,6])
      h=1.0_WP/real(n,WP)
      c=h   ! vol/h^2 on a uniform grid
      us%opr=0.0_WP
      do k=cfg%kmin_,cfg%kmax_
         do j=cfg%jmin_,cfg%jmax_
            do i=cfg%imin_,cfg%imax_
               if (solid(i,j,k,n)) cycle
               do st=1,6
                  if (.not.solid(modulo(i+off(1,st)-1,n)+1,modulo(j+off(2,st)-1,n)+1,modulo(k+off(3,st)-1,n)+1,n)) us%opr(st+1,i,j,k)=-c
               end do
               us%opr(1,i,j,k)=-sum(us%opr(2:7,i,j,k))
            end do
         end do
      end do
      call us%init()
      us%maxit=300
      us%rcvg=1.0e-10_WP
      call us%setup()
      allocate(pstar(cfg%imino_:cfg%imaxo_,cfg%jmino_:cfg%jmaxo_,cfg%kmino_:cfg%kmaxo_)); pstar=0.0_WP
      allocate(Ap,mold=pstar); Ap=0.0_WP
      do k=cfg%kmin_,cfg%kmax_
         z=(real(k,WP)-0.5_WP)*h
         do j=cfg%jmin_,cfg%jmax_
            y=(real(j,WP)-0.5_WP)*h
            do i=cfg%imin_,cfg%imax_
               x=(real(i,WP)-0.5_WP)*h
               if (.not.solid(i,j,k,n)) pstar(i,j,k)=sin(2.0_WP*pi*x)*cos(4.0_WP*pi*y)+0.5_WP*cos(2.0_WP*pi*z)
            end do
         end do
      end do
      call cfg%sync(pstar)
      do k=cfg%kmin_,cfg%kmax_
         do j=cfg%jmin_,cfg%jmax_
            do i=cfg%imin_,cfg%imax_
               Ap(i,j,k)=0.0_WP
               do st=1,7
                  Ap(i,j,k)=Ap(i,j,k)+us%opr(st,i,j,k)*pstar(i+us%stc(st,1),j+us%stc(st,2),k+us%stc(st,3))
               end do
            end do
         end do
      end do
      us%rhs=Ap
      us%sol=0.0_WP
      do k=cfg%kmin_,cfg%kmax_
         do j=cfg%jmin_,cfg%jmax_
            do i=cfg%imin_,cfg%imax_
               if (solid(i,j,k,n)) us%sol(i,j,k)=7.0_WP
            end do
         end do
      end do
      call us%solve()
      nf=0; m1=0.0_WP; m2=0.0_WP
      do k=cfg%kmin_,cfg%kmax_
         do j=cfg%jmin_,cfg%jmax_
            do i=cfg%imin_,cfg%imax_
               if (solid(i,j,k,n)) then
                  if (us%sol(i,j,k).ne.7.0_WP) nfail=nfail+1
               else
                  nf=nf+1; m1=m1+us%sol(i,j,k); m2=m2+pstar(i,j,k)
               end if
            end do
         end do
      end do
      m1=m1/real(nf,WP); m2=m2/real(nf,WP)
      err=0.0_WP; nrm=0.0_WP
      do k=cfg%kmin_,cfg%kmax_
         do j=cfg%jmin_,cfg%jmax_
            do i=cfg%imin_,cfg%imax_
               if (solid(i,j,k,n)) cycle
               err=err+(us%sol(i,j,k)-m1-pstar(i,j,k)+m2)**2
               nrm=nrm+(pstar(i,j,k)-m2)**2
            end do
         end do
      end do
      err=sqrt(err/nrm)
      call us%print_short()
      write(*,'("method ",i0," cells ",i0," it=",i0," rerr=",es10.3," err=",es10.3)') method,us%ncell,us%it,us%rerr,err
      if (err.gt.1.0e-8_WP.or.us%rerr.ge.1.0e-10_WP.or.us%ncell.ne.nf) nfail=nfail+1
      if (abs(us%sol(cfg%imin_-1,1,1)-us%sol(cfg%imax_,1,1)).gt.0.0_WP) nfail=nfail+1
      if (us%ind(cfg%imin_,cfg%jmin_,cfg%kmin_).ne.0) nfail=nfail+1
      call us%destroy()
      deallocate(pstar,Ap)
   end subroutine run_case

end program ngap_uns_demo
