! iso_c_binding interfaces of libparaht_hip.so (the C declarations are in
! include/paraht.h).  Integers are C ints passed by value, matrices are
! column-major double precision arrays passed by address.
module paraht_capi
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: PARAHT_TIME_COMPONENTS = 1
   integer(c_int), parameter :: PARAHT_STAGE2_GIVENS = 2
   integer(c_int), parameter :: PARAHT_QZ_SCHUR = 8

   interface
      integer(c_int) function paraht_ab2rht(n, A, lda, B, ldb, Q, ldq, Z, ldz, nb, p, flags) &
            bind(C, name='paraht_ab2rht')
         import :: c_int, c_double
         integer(c_int), value :: n, lda, ldb, ldq, ldz, nb, p, flags
         real(c_double) :: A(*), B(*), Q(*), Z(*)
      end function

      integer(c_int) function paraht_rht2ht(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq, flags) &
            bind(C, name='paraht_rht2ht')
         import :: c_int, c_double
         integer(c_int), value :: n, lda, ldb, ldq, ldz, r, nq, flags
         real(c_double) :: A(*), B(*), Q(*), Z(*)
      end function

      integer(c_int) function paraht_rht2ht_unblocked(n, A, lda, B, ldb, Q, ldq, Z, ldz, r) &
            bind(C, name='paraht_rht2ht_unblocked')
         import :: c_int, c_double
         integer(c_int), value :: n, lda, ldb, ldq, ldz, r
         real(c_double) :: A(*), B(*), Q(*), Z(*)
      end function

      integer(c_int) function paraht_qr_b(n, A, lda, B, ldb, Q, ldq, nb, p) &
            bind(C, name='paraht_qr_b')
         import :: c_int, c_double
         integer(c_int), value :: n, lda, ldb, ldq, nb, p
         real(c_double) :: A(*), B(*), Q(*)
      end function

      ! QZ iteration on the reduced pencil (DHGEQZ's place after the HT
      ! form); flags: PARAHT_QZ_SCHUR for the Schur form with Q, Z
      integer(c_int) function paraht_hgeqz(n, H, ldh, T, ldt, Q, ldq, Z, ldz, alphar, alphai, &
            beta, flags, info) bind(C, name='paraht_hgeqz')
         import :: c_int, c_double
         integer(c_int), value :: n, ldh, ldt, ldq, ldz, flags
         real(c_double) :: H(*), T(*), Q(*), Z(*), alphar(*), alphai(*), beta(*)
         integer(c_int) :: info
      end function

      integer(c_int) function paraht_version() bind(C, name='paraht_version')
         import :: c_int
      end function

      type(c_ptr) function paraht_status_string(status) bind(C, name='paraht_status_string')
         import :: c_int, c_ptr
         integer(c_int), value :: status
      end function

      integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
         import :: c_ptr, c_size_t
         type(c_ptr), value :: s
      end function
   end interface

contains

   ! Stops the program with the library's message when status is not 0
   ! (negative: the -status-th argument was invalid; positive: a device
   ! failure or an unsupported parameter combination).
   subroutine paraht_check(status, what)
      integer(c_int), intent(in) :: status
      character(len=*), intent(in) :: what
      type(c_ptr) :: msg
      character(kind=c_char), pointer :: chars(:)
      character(len=256) :: text
      integer :: i, nc

      if (status == 0) return
      msg = paraht_status_string(status)
      text = ' '
      if (c_associated(msg)) then
         nc = int(min(c_strlen(msg), int(len(text), c_size_t)))
         call c_f_pointer(msg, chars, [nc])
         do i = 1, nc
            text(i:i) = chars(i)
         end do
      end if
      write (*, '(A, A, I0, A, A)') what, ': paraht status ', status, ': ', trim(text)
      error stop 1
   end subroutine

   integer(c_int) function time_flag(time_components)
      logical, intent(in) :: time_components
      time_flag = 0
      if (time_components) time_flag = PARAHT_TIME_COMPONENTS
   end function

end module
