! gems_tb_hip.f90 -- the SMA-TB force call of a GEMS build, on the MI355X
! library.  GEMS keeps its own script interpreter, integrators, outputs and
! atom lists; when the environment selects the device backend
! (GEMS_BACKEND=hip), smatb_interact (src/TB.F90:457-586) hands each force
! evaluation to tb_hip_interact, which
!   * packs the group's atom positions ig%at(i)%o%pos (Neighbour.F90:114),
!   * on the first call (or when the atoms or the box change) creates a
!     device context with the group's internal types (ig%z, TB.F90:193-215),
!     the smatb parameter matrices, the box (box, Program_Types.F90:69) and
!     the skin (nb_dcut, Neighbour.F90:237),
!   * afterwards only sends the positions (gh_set_positions: the device keeps
!     its lists and decides rebuilds itself),
!   * calls gh_compute and accumulates forces (ui/A) and per-atom energies
!     (ui) into the atoms as smatb_interact does (TB.F90:534-578), and sets
!     ig%epot (ui).
! GEMS_HIP_PRECISION=mixed selects the mixed-precision form (default FP64).
! Only the minimum-image mode (mic, Program_Types.F90:98) is supported: the
! device builds its own periodic images.
module gems_tb_hip
use iso_c_binding
use gems_hip
use gems_constants, only: dp, dm
use gems_program_types, only: box, mic
use gems_neighbour, only: intergroup, nb_dcut
use gems_errors, only: werr
implicit none
private
public :: tb_hip_on, tb_hip_interact

type(c_ptr), save :: ctx = c_null_ptr
integer, save :: nat = -1
real(dp), save :: box_set(dm) = -1.0_dp
logical, save :: checked = .false., on = .false.
real(c_double), allocatable, save :: x(:,:), F(:,:), e(:)

contains

logical function tb_hip_on()
  character(len=32) :: v
  integer :: l, st
  if (.not. checked) then
    call get_environment_variable('GEMS_BACKEND', v, l, st)
    on = (st == 0 .and. l == 3)
    if (on) on = (v(1:3) == 'hip')
    checked = .true.
  endif
  tb_hip_on = on
end function

subroutine check(rc)
  integer(c_int), intent(in) :: rc
  if (rc /= 0) call werr('gems_hip: '//gh_last_error(ctx), .true.)
end subroutine

subroutine tb_hip_interact(ig, nz, zmap, ca, eps, p, q, r0, rci, rce)
  class(intergroup), intent(inout) :: ig
  integer, intent(in) :: nz, zmap(:)
  real(dp), intent(in) :: ca(nz,nz), eps(nz,nz), p(nz,nz), q(nz,nz), r0(nz,nz), rci(nz,nz), &
                          rce(nz,nz)
  integer :: i, n, prec, l, st
  integer(c_int) :: pbc(3)
  integer(c_int), allocatable :: t(:)
  real(c_double) :: et, lbox(3)
  character(len=32) :: v

  call werr('gems_hip: only the minimum-image mode (mic) is supported', .not. mic)
  n = ig%n(1)
  if (.not. allocated(x) .or. n /= nat) then
    if (allocated(x)) deallocate(x, F, e)
    allocate(x(3, n), F(3, n), e(n))
  endif
  do i = 1, n
    x(:, i) = ig%at(i)%o%pos(1:3)
  enddo

  if (.not. c_associated(ctx) .or. n /= nat .or. any(box(1:3) /= box_set(1:3))) then
    if (c_associated(ctx)) call gh_destroy(ctx)
    prec = GH_FP64
    call get_environment_variable('GEMS_HIP_PRECISION', v, l, st)
    if (st == 0 .and. l == 5) then
      if (v(1:5) == 'mixed') prec = GH_MIXED
    endif
    call check(gh_create(ctx, 0_c_int, int(prec, c_int)))
    pbc = 0
    do i = 1, 3
      if (ig%at(1)%o%pbc(i)) pbc(i) = 1
    enddo
    lbox = box(1:3)
    if (any(pbc /= 0)) then
      call check(gh_set_box(ctx, lbox, pbc))
    else
      call check(gh_set_box_free(ctx, c_null_ptr, c_null_ptr))
    endif
    allocate(t(n))
    do i = 1, n
      t(i) = zmap(ig%at(i)%o%z) - 1
    enddo
    call check(gh_set_atoms(ctx, int(n, c_int), x, t))
    call check(gh_set_smatb(ctx, int(nz, c_int), ca, eps, p, q, r0, rci, rce))
    call check(gh_set_skin(ctx, real(nb_dcut, c_double)))
    nat = n
    box_set(1:3) = box(1:3)
  else
    call check(gh_set_positions(ctx, x))
  endif

  call check(gh_compute(ctx, F, e, et))
  do i = 1, n
    ig%at(i)%o%force(1:3) = ig%at(i)%o%force(1:3) + F(:, i)
    ig%at(i)%o%epot = ig%at(i)%o%epot + e(i)
  enddo
  ig%epot = et
end subroutine

end module gems_tb_hip
