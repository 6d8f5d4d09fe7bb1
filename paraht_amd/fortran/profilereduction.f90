! Benchmark driver with the command line and result line of the reference's
! src/profilereduction.f90:
!   profilereduction n algorithm [r p nq [time_components]] [pencil=FILE] [out=FILE] [qz=1]
! algorithm 1: ab2rht_simple + rht2ht_householder (modules ab2rht / rht2ht,
! here the GPU drop-ins); algorithm 2: LAPACK DGGHD3 on the host.  The pencil
! is A uniform on [0,1), B the R factor of a uniform [0,1) matrix, Q = Z = I,
! from the compiler's random_number seeded with 1302, or read from a pencil
! file (pencil=FILE; n = 0 takes the file's order).  A file's B need not be
! triangular: if it has nonzeros below the diagonal, its QR comes first,
! carried into A and Q (on the GPU, paraht_qr_b, for algorithm 1; LAPACK
! dgeqrf/dormqr/dorgqr for algorithm 2) and timed with the reduction, and the
! errors are those against the pencil in the file.  out=FILE writes the
! reduced H, T, Q, Z.  qz=1 then runs the QZ iteration on the reduced pencil
! (eigenvalues only: paraht_hgeqz on the GPU for algorithm 1, LAPACK DHGEQZ
! for algorithm 2; the reference stops at the HT form) and prints a second
! line with its time, DHGEQZ's info and the number of infinite eigenvalues,
! for algorithm 1 also the largest chordal distance from a GPU eigenvalue to
! the nearest of DHGEQZ's on the same pencil.  Prints
!   n: ;runtime: ;error: ||Q^T A0 Z - H||/||A0|| ||Q^T B0 Z - T||/||B0||
! (and for algorithm 1 the two phase times and r, p, nq).
!
! Pencil files (paraht_amd.pencils.save_pencil / load_pencil): the 8 bytes
! "PHTPENC1", the order n as a little-endian 64-bit integer, then A and B,
! each n x n column-major little-endian FP64.  Output files: "PHTREDU1", n,
! then H, T, Q, Z in the same layout.
program profilereduction
   use ab2rht
   use rht2ht
   use paraht_capi, only: paraht_qr_b, paraht_hgeqz, paraht_check
   implicit none

   integer :: n, algorithm, r, p, nq, flagarg, info, nseed, k, npos, iarg
   logical :: time_components, want_qz
   character(len=4096) :: pencil_file, out_file, full
   double precision, allocatable :: A(:, :), B(:, :), Q(:, :), Z(:, :), A0(:, :), B0(:, :), W(:, :)
   double precision, allocatable :: tau(:), work(:)
   integer, allocatable :: seed(:)
   double precision :: t1, t2, tqr, total, ea, eb
   logical :: general
   integer(8) :: c0, c1, rate
   character(len=32) :: arg
   double precision, external :: dlange

   r = 0; p = 0; nq = 0
   time_components = .false.
   pencil_file = ''
   out_file = ''
   want_qz = .false.
   ! key=value options may follow the positional arguments
   npos = 0
   do iarg = 1, command_argument_count()
      call get_command_argument(iarg, full)
      if (full(1:7) == 'pencil=') then
         pencil_file = full(8:)
      else if (full(1:4) == 'out=') then
         out_file = full(5:)
      else if (full(1:3) == 'qz=') then
         want_qz = full(4:4) == '1'
      else
         npos = npos + 1
      end if
   end do
   if (npos < 2) then
      write (*, *) 'usage: profilereduction n algorithm (r p nq time_components) (pencil=FILE) (out=FILE) (qz=1)'
      stop
   end if
   call get_command_argument(1, arg); read (arg, *) n
   call get_command_argument(2, arg); read (arg, *) algorithm
   if (algorithm == 1) then
      if (npos < 5) then
         write (*, *) 'algorithm 1 needs the parameters r, p and nq'
         stop
      end if
      call get_command_argument(3, arg); read (arg, *) r
      call get_command_argument(4, arg); read (arg, *) p
      call get_command_argument(5, arg); read (arg, *) nq
      if (npos >= 6) then
         call get_command_argument(6, arg); read (arg, *) flagarg
         time_components = flagarg == 1
      end if
   end if
   if (len_trim(pencil_file) > 0) n = pencil_order(trim(pencil_file), n)

   allocate (A(n, n), B(n, n), Q(n, n), Z(n, n), A0(n, n), B0(n, n), W(n, n))
   allocate (tau(n), work(max(1, 64*n)))
   if (len_trim(pencil_file) > 0) then
      call read_pencil(trim(pencil_file), n, A, B)
   else
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
   end if
   call identity(Q)
   call identity(Z)
   A0 = A
   B0 = B

   general = .false.
   do k = 1, n - 1
      if (any(B(k + 1:n, k) /= 0.0d0)) general = .true.
   end do
   tqr = 0.0d0
   if (general) then
      call system_clock(c0)
      if (algorithm == 1) then
         call paraht_check(paraht_qr_b(n, A, n, B, n, Q, n, 64, 2), 'paraht_qr_b')
      else
         call dgeqrf(n, n, B, n, tau, work, size(work), info)
         call dormqr('L', 'T', n, n, n, B, n, tau, A, n, work, size(work), info)
         Q = B
         call dorgqr(n, n, n, Q, n, tau, work, size(work), info)
         do k = 1, n - 1
            B(k + 1:n, k) = 0.0d0
         end do
      end if
      call system_clock(c1, rate)
      tqr = dble(c1 - c0)/dble(rate)
   end if

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
      total = tqr + t1 + t2
   else
      call system_clock(c0)
      call dgghd3('V', 'V', n, 1, n, A, n, B, n, Q, n, Z, n, W, n*n, info)
      call system_clock(c1, rate)
      total = tqr + dble(c1 - c0)/dble(rate)
   end if

   ea = residual(A0, A)
   eb = residual(B0, B)
   if (len_trim(out_file) > 0) call write_reduced(trim(out_file), n, A, B, Q, Z)
   if (algorithm == 1) then
      write (*, '("n:",I6," ;runtime:",F16.6," ;error:",2ES16.6, " ;phase 1 time:",F16.6, &
             &" ;phase 2 time:",F16.6," ;r:",I4," ;p:",I4," ;nq:",I4)') n, total, ea, eb, t1, t2, r, p, nq
   else
      write (*, '("n:",I6," ;runtime:",F16.6," ;error:",2ES16.6)') n, total, ea, eb
   end if
   if (want_qz) call run_qz()

contains

   ! The QZ iteration on the reduced pencil (A, B now hold H, T).
   subroutine run_qz()
      double precision, allocatable :: H(:, :), T(:, :), ar(:), ai(:), be(:)
      double precision, allocatable :: lr(:), li(:), lb(:), wk(:)
      double precision :: dq(1, 1), tq, dmax, dmin
      integer :: st, qinfo, linfo, i, j, ninf
      allocate (H(n, n), T(n, n), ar(n), ai(n), be(n), wk(max(1, n)))
      H = A
      T = B
      call system_clock(c0)
      if (algorithm == 1) then
         st = paraht_hgeqz(n, H, n, T, n, dq, 1, dq, 1, ar, ai, be, 0, qinfo)
         call paraht_check(st, 'paraht_hgeqz')
      else
         call dhgeqz('E', 'N', 'N', n, 1, n, H, n, T, n, ar, ai, be, dq, 1, dq, 1, wk, n, qinfo)
      end if
      call system_clock(c1, rate)
      tq = dble(c1 - c0)/dble(rate)
      ninf = count(be == 0.0d0)
      if (algorithm == 1) then
         allocate (lr(n), li(n), lb(n))
         H = A
         T = B
         call dhgeqz('E', 'N', 'N', n, 1, n, H, n, T, n, lr, li, lb, dq, 1, dq, 1, wk, n, linfo)
         dmax = 0.0d0
         do i = 1, n
            dmin = huge(1.0d0)
            do j = 1, n
               dmin = min(dmin, chordal(ar(i), ai(i), be(i), lr(j), li(j), lb(j)))
            end do
            dmax = max(dmax, dmin)
         end do
         write (*, '("qz: time:",F16.6," ;info:",I6," ;infinite:",I6," ;chordal vs dhgeqz:",ES16.6)') &
            tq, qinfo, ninf, dmax
      else
         write (*, '("qz: time:",F16.6," ;info:",I6," ;infinite:",I6)') tq, qinfo, ninf
      end if
   end subroutine

   ! chordal distance between (a1 + i c1)/b1 and (a2 + i c2)/b2
   double precision function chordal(a1, c1, b1, a2, c2, b2)
      double precision, intent(in) :: a1, c1, b1, a2, c2, b2
      complex(8) :: x, y
      x = cmplx(a1, c1, 8)
      y = cmplx(a2, c2, 8)
      chordal = abs(x*b2 - b1*y)/(sqrt(abs(x)**2 + b1**2)*sqrt(abs(y)**2 + b2**2))
   end function

   ! The order of a pencil file (checked against a nonzero n from the
   ! command line).
   integer function pencil_order(path, n_arg)
      character(len=*), intent(in) :: path
      integer, intent(in) :: n_arg
      character(len=8) :: magic
      integer(8) :: n8
      integer :: u
      open (newunit=u, file=path, access='stream', form='unformatted', status='old', action='read')
      read (u) magic, n8
      close (u)
      if (magic /= 'PHTPENC1') error stop 'profilereduction: not a pencil file'
      if (n_arg /= 0 .and. n_arg /= int(n8)) error stop 'profilereduction: n differs from the file'
      pencil_order = int(n8)
   end function

   subroutine read_pencil(path, n, A, B)
      character(len=*), intent(in) :: path
      integer, intent(in) :: n
      double precision, intent(out) :: A(n, n), B(n, n)
      character(len=8) :: magic
      integer(8) :: n8
      integer :: u
      open (newunit=u, file=path, access='stream', form='unformatted', status='old', action='read')
      read (u) magic, n8, A, B
      close (u)
   end subroutine

   subroutine write_reduced(path, n, H, T, Q, Z)
      character(len=*), intent(in) :: path
      integer, intent(in) :: n
      double precision, intent(in) :: H(n, n), T(n, n), Q(n, n), Z(n, n)
      integer :: u
      open (newunit=u, file=path, access='stream', form='unformatted', status='replace', action='write')
      write (u) 'PHTREDU1', int(n, 8), H, T, Q, Z
      close (u)
   end subroutine

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
