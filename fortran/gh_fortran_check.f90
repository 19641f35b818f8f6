! Drives the library through the ISO_C_BINDING module the way a GEMS build
! would: reads a configuration (n, box, positions) from a text file, sets the
! Au SMA-TB potential, evaluates forces once (interact) and runs velocity
! Verlet steps (dinamica) after `set tempgdist` from the lcg stream.
!   gh_fortran_check <in.txt> <out.txt> <precision 0|1>
program gh_fortran_check
  use iso_c_binding
  use gems_hip
  implicit none
  type(c_ptr) :: ctx
  integer(c_int) :: n, rc, prec, i
  real(c_double) :: L(3), A(1,1), xi(1,1), p(1,1), q(1,1), r0(1,1), rci(1,1), rce(1,1)
  real(c_double), allocatable :: x(:,:), F(:,:), e(:)
  integer(c_int), allocatable :: ty(:)
  real(c_double) :: etot, en(3, 2), mass(1)
  character(len=512) :: fin, fout, sprec

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, sprec)
  read(sprec, *) prec

  open(10, file=trim(fin), action='read')
  read(10, *) n
  read(10, *) L
  allocate(x(3, n), F(3, n), e(n), ty(n))
  do i = 1, n
    read(10, *) x(:, i)
  enddo
  close(10)
  ty = 0

  rc = gh_create(ctx, 0, prec);                       call check('create')
  rc = gh_set_box(ctx, L, [1_c_int, 1_c_int, 1_c_int]); call check('box')
  rc = gh_set_atoms(ctx, n, x, ty);                   call check('atoms')
  mass(1) = real(196.97, c_double)                    ! Elements.F90:142 (default real)
  rc = gh_set_masses(ctx, 1, mass);                   call check('masses')
  A = 0.2061_c_double; xi = 1.790_c_double; p = 10.229_c_double; q = 4.036_c_double
  r0 = 2.884_c_double; rci = 4.078_c_double; rce = 4.078_c_double * sqrt(1.5_c_double)
  rc = gh_set_smatb(ctx, 1, A, xi, p, q, r0, rci, rce); call check('smatb')
  rc = gh_compute(ctx, F, e, etot);                   call check('compute')
  rc = gh_prng_lcg(ctx, 4242);                        call check('prng')
  rc = gh_set_gdist(ctx, 300.0_c_double);             call check('gdist')
  rc = gh_integrator_vv(ctx, 0.002_c_double);         call check('vv')
  rc = gh_run(ctx, 20, 10, en);                       call check('run')

  open(11, file=trim(fout), action='write')
  write(11, '(es25.17)') etot
  do i = 1, n
    write(11, '(3es25.17)') F(:, i)
  enddo
  write(11, '(3es25.17)') en
  close(11)
  call gh_destroy(ctx)

contains
  subroutine check(what)
    character(*), intent(in) :: what
    if (rc /= 0) then
      write(*, *) 'gems_hip error in ', what, ': ', gh_last_error(ctx)
      stop 1
    endif
  end subroutine
end program
