!> Self-checking Fortran driver of the ngap_str drop-in, written the way an
!> NGA2 driver uses hypre_str (examples/channel/src/simulation.f90:137-140,
!> 315-321): construct, set stc, init, fill opr, setup, fill rhs/sol, solve.
!> A manufactured smooth solution p* gives rhs = A p*; the solve must return
!> p* (mean-free) to 1e-8 and reach the requested residual, in FP64 and in
!> mixed precision, on a periodic box and on a channel with walls in y.
!> Exit status 0 when every case passes.
program ngap_demo
   use precision,      only: WP
   use config_class,   only: config
   use ngap_str_class
   implicit none
   integer :: n,nfail
   character(len=32) :: arg
   n=32
   if (command_argument_count().ge.1) then
      call get_command_argument(1,arg); read(arg,*) n
   end if
   nfail=0
   call run_case('periodic fp64 ',.true. ,ngap_fp64 ,nfail)
   call run_case('periodic mixed',.true. ,ngap_mixed,nfail)
   call run_case('channel  fp64 ',.false.,ngap_fp64 ,nfail)
   call run_case('channel  mixed',.false.,ngap_mixed,nfail)
   if (nfail.gt.0) then
      write(*,'("NGAP_DEMO FAIL ",i0)') nfail
      error stop 1
   end if
   write(*,'("NGAP_DEMO OK")')

contains

   subroutine run_case(label,yper,prec,nfail)
      character(len=*), intent(in) :: label
      logical, intent(in) :: yper
      integer, intent(in) :: prec
      integer, intent(inout) :: nfail
      type(config), target :: cfg
      type(ngap_str) :: ps
      real(WP), dimension(:,:,:), allocatable :: pstar,Ap
      real(WP) :: vol,cx,cy,cz,pi,x,y,z,err,nrm,m1,m2
      integer :: i,j,k,st
      pi=4.0_WP*atan(1.0_WP)
      cfg=config(n,n,n,1,1.0_WP,1.0_WP,1.0_WP,.true.,yper,.true.)
      cfg%name='demo'
      ps=ngap_str(cfg=cfg,name='Pressure',method=pcg_pfmg,nst=7)
      ps%maxlevel=0
      ps%maxit=200
      ps%rcvg=1.0e-10_WP
      ps%precision=prec
      ! 7-point stencil map (incomp_class.f90:758-764)
      ps%stc(1,:)=[ 0, 0, 0]
      ps%stc(2,:)=[+1, 0, 0]
      ps%stc(3,:)=[-1, 0, 0]
      ps%stc(4,:)=[ 0,+1, 0]
      ps%stc(5,:)=[ 0,-1, 0]
      ps%stc(6,:)=[ 0, 0,+1]
      ps%stc(7,:)=[ 0, 0,-1]
      call ps%init()
      ! -vol*div(grad) on the uniform mesh; walls in y when not periodic
      vol=cfg%dx*cfg%dy*cfg%dz
      cx=vol/cfg%dx**2; cy=vol/cfg%dy**2; cz=vol/cfg%dz**2
      do k=cfg%kmin_,cfg%kmax_
         do j=cfg%jmin_,cfg%jmax_
            do i=cfg%imin_,cfg%imax_
               ps%opr(2:3,i,j,k)=-cx
               ps%opr(4:5,i,j,k)=-cy
               ps%opr(6:7,i,j,k)=-cz
               if (.not.yper.and.j.eq.cfg%jmin_) ps%opr(5,i,j,k)=0.0_WP
               if (.not.yper.and.j.eq.cfg%jmax_) ps%opr(4,i,j,k)=0.0_WP
               ps%opr(1,i,j,k)=-sum(ps%opr(2:7,i,j,k))
            end do
         end do
      end do
      call ps%setup()
      ! manufactured solution and rhs = A p*
      allocate(pstar(cfg%imino_:cfg%imaxo_,cfg%jmino_:cfg%jmaxo_,cfg%kmino_:cfg%kmaxo_)); pstar=0.0_WP
      allocate(Ap,mold=pstar); Ap=0.0_WP
      do k=cfg%kmin_,cfg%kmax_
         z=(real(k,WP)-0.5_WP)*cfg%dz
         do j=cfg%jmin_,cfg%jmax_
            y=(real(j,WP)-0.5_WP)*cfg%dy
            do i=cfg%imin_,cfg%imax_
               x=(real(i,WP)-0.5_WP)*cfg%dx
               pstar(i,j,k)=sin(2.0_WP*pi*x)*cos(pi*y)+0.5_WP*cos(2.0_WP*pi*z)
            end do
         end do
      end do
      call cfg%sync(pstar)
      do k=cfg%kmin_,cfg%kmax_
         do j=cfg%jmin_,cfg%jmax_
            do i=cfg%imin_,cfg%imax_
               Ap(i,j,k)=0.0_WP
               do st=1,7
                  Ap(i,j,k)=Ap(i,j,k)+ps%opr(st,i,j,k)*pstar(i+ps%stc(st,1),j+ps%stc(st,2),k+ps%stc(st,3))
               end do
            end do
         end do
      end do
      ps%rhs=Ap
      ps%sol=0.0_WP
      call ps%solve()
      ! compare mean-free
      m1=sum(ps%sol(cfg%imin_:cfg%imax_,cfg%jmin_:cfg%jmax_,cfg%kmin_:cfg%kmax_))/real(n,WP)**3
      m2=sum(pstar(cfg%imin_:cfg%imax_,cfg%jmin_:cfg%jmax_,cfg%kmin_:cfg%kmax_))/real(n,WP)**3
      err=sqrt(sum((ps%sol(cfg%imin_:cfg%imax_,cfg%jmin_:cfg%jmax_,cfg%kmin_:cfg%kmax_)-m1 &
      &            -pstar(cfg%imin_:cfg%imax_,cfg%jmin_:cfg%jmax_,cfg%kmin_:cfg%kmax_)+m2)**2))
      nrm=sqrt(sum((pstar(cfg%imin_:cfg%imax_,cfg%jmin_:cfg%jmax_,cfg%kmin_:cfg%kmax_)-m2)**2))
      err=err/nrm
      call ps%print_short()
      write(*,'(a," it=",i0," rerr=",es10.3," true=",es10.3," err=",es10.3," levels=",i0)') label,ps%it,ps%rerr, &
      &     ps%stats%true_rerr,err,ps%stats%nlevels
      if (err.gt.1.0e-8_WP.or.ps%stats%true_rerr.gt.1.0e-10_WP.or.ps%it.le.0) nfail=nfail+1
      ! ghosts synced like cfg%sync after the solve
      if (yper.and.abs(ps%sol(cfg%imin_-1,1,1)-ps%sol(cfg%imax_,1,1)).gt.0.0_WP) nfail=nfail+1
      call ps%destroy()
      deallocate(pstar,Ap)
   end subroutine run_case

end program ngap_demo
