! Benchmark driver with the command line and result line of the reference's
! src/profilereduction.f90:
!   profilereduction n algorithm [r p nq [time_components]]
! algorithm 1: ab2rht_simple + rht2ht_householder (modules ab2rht / rht2ht,
! here the GPU drop-ins); algorithm 2: LAPACK DGGHD3 on the host.  The pencil
! is A uniform on [0,1), B the R factor of a uniform [0,1) matrix, Q = Z = I,
! from the compiler's random_number seeded with 1302.  Prints
!   n: ;runtime: ;error: ||Q^T A0 Z - H||/||A0|| ||Q^T B0 Z - T||/||B0||
! (and for algorithm 1 the two phase times and r, p, nq).
program profilereduction
   use ab2rht
   use rht2ht
   implicit none

   integer :: n, algorithm, r, p, nq, flagarg, info, nseed, k
   logical :: time_components
   double precision, allocatable :: A(:, :), B(:, :), Q(:, :), Z(:, :), A0(:, :), B0(:, :), W(:, :)
   double precision, allocatable :: tau(:), work(:)
   integer, allocatable :: seed(:)
   double precision :: t1, t2, total, ea, eb
   integer(8) :: c0, c1, rate
   character(len=32) :: arg
   double precision, external :: dlange

   r = 0; p = 0; nq = 0
   time_components = .false.
   if (command_argument_count() < 2) then
      write (*, *) 'usage: profilereduction n algorithm (r p nq time_components)'
      stop
   end if
   call get_command_argument(1, arg); read (arg, *) n
   call get_command_argument(2, arg); read (arg, *) algorithm
   if (algorithm == 1) then
      if (command_argument_count() < 5) then
         write (*, *) 'algorithm 1 needs the parameters r, p and nq'
         stop
      end if
      call get_command_argument(3, arg); read (arg, *) r
      call get_command_argument(4, arg); read (arg, *) p
      call get_command_argument(5, arg); read (arg, *) nq
      if (command_argument_count() >= 6) then
         call get_command_argument(6, arg); read (arg, *) flagarg
         time_components = flagarg == 1
      end if
   end if

   allocate (A(n, n), B(n, n), Q(n, n), Z(n, n), A0(n, n), B0(n, n), W(n, n))
   allocate (tau(n), work(max(1, 64*n)))
   call random_seed(size=nseed)
   allocate (seed(nseed))
   seed = 0
   seed(1) = 1302
   call random_seed(put=seed)
   call random_number(A)
   call random_number(B)
   call dgeqrf(n, n, B, n, tau, work, size(work), info)
   do k = 1, n - 1
      B(k + 1:n, k) = 0.0d0
   end do
   call identity(Q)
   call identity(Z)
   A0 = A
   B0 = B

   t1 = 0.0d0; t2 = 0.0d0
   if (algorithm == 1) then
      call system_clock(c0)
      call ab2rht_simple(n, A, n, B, n, Q, n, Z, n, r, p, time_components)
      call system_clock(c1, rate)
      t1 = dble(c1 - c0)/dble(rate)
      call system_clock(c0)
      call rht2ht_householder(n, A, n, B, n, Q, n, Z, n, r, nq, time_components)
      call system_clock(c1, rate)
      t2 = dble(c1 - c0)/dble(rate)
      total = t1 + t2
   else
      call system_clock(c0)
      call dgghd3('V', 'V', n, 1, n, A, n, B, n, Q, n, Z, n, W, n*n, info)
      call system_clock(c1, rate)
      total = dble(c1 - c0)/dble(rate)
   end if

   ea = residual(A0, A)
   eb = residual(B0, B)
   if (algorithm == 1) then
      write (*, '("n:",I6," ;runtime:",F16.6," ;error:",2ES16.6, " ;phase 1 time:",F16.6, &
             &" ;phase 2 time:",F16.6," ;r:",I4," ;p:",I4," ;nq:",I4)') n, total, ea, eb, t1, t2, r, p, nq
   else
      write (*, '("n:",I6," ;runtime:",F16.6," ;error:",2ES16.6)') n, total, ea, eb
   end if

contains

   subroutine identity(M)
      double precision, intent(out) :: M(:, :)
      integer :: i
      M = 0.0d0
      do i = 1, size(M, 1)
         M(i, i) = 1.0d0
      end do
   end subroutine

   ! ||Q^T M0 Z - M||_F / ||M0||_F
   double precision function residual(M0, M)
      double precision, intent(in) :: M0(:, :), M(:, :)
      double precision, allocatable :: D(:, :)
      allocate (D(n, n))
      call dgemm('T', 'N', n, n, n, 1.0d0, Q, n, M0, n, 0.0d0, W, n)
      D = M
      call dgemm('N', 'N', n, n, n, 1.0d0, W, n, Z, n, -1.0d0, D, n)
      residual = dlange('F', n, n, D, n, work)/dlange('F', n, n, M0, n, work)
   end function

end program
