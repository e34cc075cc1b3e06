! Test doubles of the two reference modules that the minors shims
! (gemini_unified_amd/fortran/geminiMinorsShim.f90) use, reduced to what the
! shims and the test program touch: gemini_integration_environment with omre,
! omim and dll1 (geminiIntegrationEnvironment.f90) and error_message with add
! and .level. (errorMessage.f90).  A host build uses the reference's own modules.
module geminiIntegrationEnvironment
   implicit none
   type gemini_integration_environment
      double precision :: omre = 0.d0, omim = 0.d0, dll1 = 0.d0
   end type
end module geminiIntegrationEnvironment

module errorMessage
   implicit none
   type error_message
      integer :: level = 0
      character(len=256) :: message = ''
      character(len=64) :: fctname = ''
   end type
   interface add; module procedure addErrorMessage; end interface
   interface operator (.level.); module procedure levelErrorMessage; end interface
contains
   subroutine addErrorMessage(this, level, message, fctname)
      type (error_message) :: this
      integer :: level
      character(len=*) :: message, fctname
      this%level = max(this%level, level)
      this%message = message
      this%fctname = fctname
   end subroutine addErrorMessage
   integer function levelErrorMessage(this)
      type (error_message), intent(in) :: this
      levelErrorMessage = this%level
   end function levelErrorMessage
end module errorMessage
