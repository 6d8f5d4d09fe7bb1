! Drop-in for the reference's module rht2ht (src/rht2ht.f90): the three
! stage-2 entry points with their names and argument lists
! (src/rht2ht.f90:696-700, 1346-1351, 2053-2057).  rht2ht_householder and
! rht2ht_householder_unblocked run the GPU's Householder bulge chase (the
! reference's reflectors, in its order), rht2ht_givens the GPU's Givens chase,
! which reaches the same Hessenberg-triangular pencil up to the signs of rows
! and columns.
module rht2ht
   use paraht_capi
   implicit none

contains

   subroutine rht2ht_householder(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq, time_components)
      integer, intent(in) :: n, lda, ldb, ldq, ldz, r, nq
      double precision, intent(inout) :: A(lda, *), B(ldb, *), Z(ldz, *), Q(ldq, *)
      logical :: time_components

      call paraht_check(paraht_rht2ht(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq, &
                                      time_flag(time_components)), 'rht2ht_householder')
   end subroutine

   subroutine rht2ht_givens(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq)
      integer, intent(in) :: n, lda, ldb, ldq, ldz, r, nq
      double precision, intent(inout) :: A(lda, *), B(ldb, *), Z(ldz, *), Q(ldq, *)

      call paraht_check(paraht_rht2ht(n, A, lda, B, ldb, Q, ldq, Z, ldz, r, nq, &
                                      PARAHT_STAGE2_GIVENS), 'rht2ht_givens')
   end subroutine

   subroutine rht2ht_householder_unblocked(n, A, lda, B, ldb, Q, ldq, Z, ldz, r)
      integer, intent(in) :: n, lda, ldb, ldq, ldz, r
      double precision, intent(inout) :: A(lda, *), B(ldb, *), Z(ldz, *), Q(ldq, *)

      call paraht_check(paraht_rht2ht_unblocked(n, A, lda, B, ldb, Q, ldq, Z, ldz, r), &
                        'rht2ht_householder_unblocked')
   end subroutine

end module
