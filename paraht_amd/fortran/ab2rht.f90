! Drop-in for the reference's module ab2rht (src/ab2rht.f90): the same
! module and procedure name and the same argument list, so callers such as
! src/profilereduction.f90:136 compile unchanged; the reduction runs on the
! GPU through libparaht_hip.so.
module ab2rht
   use paraht_capi
   implicit none

contains

   ! (A, B) -> r-Hessenberg-triangular form with nb = r and p nb-row blocks
   ! per reflector block; Q <- Q Q1, Z <- Z Z1 (src/ab2rht.f90:8-13).
   subroutine ab2rht_simple(n, A, lda, B, ldb, Q, ldq, Z, ldz, nb, p, time_components)
      integer, intent(in) :: n, lda, ldb, ldq, ldz, nb, p
      double precision, intent(inout) :: A(lda, *), B(ldb, *), Z(ldz, *), Q(ldq, *)
      logical :: time_components

      call paraht_check(paraht_ab2rht(n, A, lda, B, ldb, Q, ldq, Z, ldz, nb, p, &
                                      time_flag(time_components)), 'ab2rht_simple')
   end subroutine

end module
