!> Test scaffolding: the linsol plugin contract (NGA2
!> src/solver/linsol_class.f90:16-71), restated so the ngap_str shim can be
!> compiled and exercised outside an NGA2 tree. Field names and the deferred
!> procedure set are the contract the shim implements.
module linsol_class
   use precision,    only: WP
   use config_class, only: config
   use string,       only: str_medium
   implicit none
   private
   public :: linsol

   type, abstract :: linsol
      type(config), pointer :: cfg
      character(len=str_medium) :: name
      integer :: method
      integer :: nst
      integer, dimension(:,:), allocatable :: stc
      integer, dimension(:,:,:), allocatable :: stmap
      real(WP), dimension(:,:,:,:), allocatable :: opr
      real(WP), dimension(:,:,:), allocatable :: rhs
      real(WP), dimension(:,:,:), allocatable :: sol
      logical :: setup_done
      integer :: maxit
      real(WP) :: rcvg
      real(WP) :: acvg
      integer :: it
      real(WP) :: rerr
      real(WP) :: aerr
   contains
      procedure(in_noarg), deferred :: print_short
      procedure(in_noarg), deferred :: print
      procedure(in_noarg), deferred :: log
      procedure(inout_noarg), deferred :: init
      procedure(inout_noarg), deferred :: setup
      procedure(inout_noarg), deferred :: solve
      procedure(inout_noarg), deferred :: destroy
   end type linsol

   abstract interface
      subroutine in_noarg(this)
         import linsol
         class(linsol), intent(in) :: this
      end subroutine in_noarg
      subroutine inout_noarg(this)
         import linsol
         class(linsol), intent(inout) :: this
      end subroutine inout_noarg
   end interface

end module linsol_class
