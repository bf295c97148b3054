!> Test scaffolding: the few NGA2 library modules the ngap_str shim uses
!> (precision, string, messager, param), restated with only the names the shim
!> needs. Inside an NGA2 build the shim compiles against NGA2's own modules
!> (src/libraries/precision_dp.f90, string.f90, messager.f90, param.f90).
module precision
   use, intrinsic :: iso_fortran_env, only: real64
   implicit none
   integer, parameter :: WP=real64
end module precision

module string
   implicit none
   integer, parameter :: str_short=8
   integer, parameter :: str_medium=64
   integer, parameter :: str_long=8192
end module string

module messager
   use, intrinsic :: iso_fortran_env, only: error_unit,output_unit
   implicit none
contains
   subroutine die(text)
      character(len=*), intent(in) :: text
      write(error_unit,'(a)') trim(text)
      error stop 2
   end subroutine die
   subroutine log(text)
      character(len=*), intent(in) :: text
      write(output_unit,'(a)') trim(text)
   end subroutine log
end module messager

module param
   implicit none
   integer :: verbose=0
end module param
