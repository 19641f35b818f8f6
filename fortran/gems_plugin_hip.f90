! gems_plugin_hip.f90 -- the force calls of a GEMS build, on the MI355X
! library.  GEMS keeps its own script interpreter, integrators, outputs and
! atom lists; when the environment selects the device backend
! (GEMS_BACKEND=hip), smatb_interact (src/TB.F90:457-586) hands each force
! evaluation to tb_hip_interact and lj_interact (src/Pairs.F90:207-279) to
! lj_hip_interact.  Each of them
!   * packs the group's atom positions ig%at(i)%o%pos (Neighbour.F90:114),
!   * on the first call (or when the atoms or the box change) creates a
!     device context with the group's types, the potential's parameters, the
!     box (box, Program_Types.F90:69) and the skin (nb_dcut, Neighbour.F90:237),
!   * afterwards only sends the positions (gh_set_positions: the device keeps
!     its lists and decides rebuilds itself),
!   * calls gh_compute and accumulates forces (ui/A) and per-atom energies
!     (ui) into the atoms as the reference routines do, and sets ig%epot (ui).
! GEMS_HIP_PRECISION=mixed selects the mixed-precision form (default FP64).
! Both periodic-image modes are served (mic, Program_Types.F90:98): `box mic
! T` pairs nearest images, `box mic F` every first-shell image as GEMS's ghost
! atoms do (gh_set_images); the device builds its own images either way and
! only the group's local atoms (ig%n(1)) travel.  One context per potential
! kind: a script with one SMA-TB group and one LJ group is served, more are
! not, and only self (`under`) groups: a cross group stops with an error.
module gems_plugin_hip
use iso_c_binding
use gems_hip
use gems_constants, only: dp, dm
use gems_program_types, only: box, mic
use gems_neighbour, only: intergroup, nb_dcut
use gems_errors, only: werr
implicit none
private
public :: gems_hip_on, tb_hip_interact, lj_hip_interact

! device state of one interaction group
type plugin
  type(c_ptr) :: ctx = c_null_ptr
  integer :: nat = -1
  real(dp) :: box_set(dm) = -1.0_dp
  logical :: mic_set = .true.
  real(c_double), allocatable :: x(:,:), F(:,:), e(:)
end type

type(plugin), save, target :: tb, lj
logical, save :: checked = .false., on = .false.

contains

logical function gems_hip_on()
  character(len=32) :: v
  integer :: l, st
  if (.not. checked) then
    call get_environment_variable('GEMS_BACKEND', v, l, st)
    on = (st == 0 .and. l == 3)
    if (on) on = (v(1:3) == 'hip')
    checked = .true.
  endif
  gems_hip_on = on
end function

subroutine check(pl, rc)
  type(plugin), intent(in) :: pl
  integer(c_int), intent(in) :: rc
  if (rc /= 0) call werr('gems_hip: '//gh_last_error(pl%ctx), .true.)
end subroutine

! Positions in; true when the context must be (re)built for this group.
logical function pack_positions(pl, ig)
  type(plugin), intent(inout) :: pl
  class(intergroup), intent(in) :: ig
  integer :: i, n
  ! cross groups (`A with B`, `A feels B`) hold two atom lists; the device
  ! form of `feels` is gh_set_pair_masks, not wired into this hook
  call werr('gems_hip: only self (`under`) interaction groups are served', .not. ig%self)
  n = ig%n(1)
  if (.not. allocated(pl%x) .or. n /= pl%nat) then
    if (allocated(pl%x)) deallocate(pl%x, pl%F, pl%e)
    allocate(pl%x(3, n), pl%F(3, n), pl%e(n))
  endif
  do i = 1, n
    pl%x(:, i) = ig%at(i)%o%pos(1:3)
  enddo
  pack_positions = .not. c_associated(pl%ctx) .or. n /= pl%nat .or. &
                   any(box(1:3) /= pl%box_set(1:3)) .or. (mic .neqv. pl%mic_set)
end function

! A new context with the group's box and atoms (types t, 0-based).
subroutine new_context(pl, ig, t)
  type(plugin), intent(inout) :: pl
  class(intergroup), intent(in) :: ig
  integer(c_int), intent(in) :: t(:)
  integer :: i, prec, l, st
  integer(c_int) :: pbc(3)
  real(c_double) :: lbox(3)
  character(len=32) :: v
  if (c_associated(pl%ctx)) call gh_destroy(pl%ctx)
  prec = GH_FP64
  call get_environment_variable('GEMS_HIP_PRECISION', v, l, st)
  if (st == 0 .and. l == 5) then
    if (v(1:5) == 'mixed') prec = GH_MIXED
  endif
  call check(pl, gh_create(pl%ctx, 0_c_int, int(prec, c_int)))
  pbc = 0
  do i = 1, 3
    if (ig%at(1)%o%pbc(i)) pbc(i) = 1
  enddo
  lbox = box(1:3)
  if (any(pbc /= 0)) then
    call check(pl, gh_set_box(pl%ctx, lbox, pbc))
  else
    call check(pl, gh_set_box_free(pl%ctx, c_null_ptr, c_null_ptr))
  endif
  call check(pl, gh_set_images(pl%ctx, merge(0_c_int, 1_c_int, mic)))
  call check(pl, gh_set_atoms(pl%ctx, int(ig%n(1), c_int), pl%x, t))
  call check(pl, gh_set_skin(pl%ctx, real(nb_dcut, c_double)))
  pl%nat = ig%n(1)
  pl%box_set(1:3) = box(1:3)
  pl%mic_set = mic
end subroutine

! Forces and energies out, accumulated as the reference routines do.
subroutine compute(pl, ig)
  type(plugin), intent(inout) :: pl
  class(intergroup), intent(inout) :: ig
  integer :: i
  real(c_double) :: et
  call check(pl, gh_compute(pl%ctx, pl%F, pl%e, et))
  do i = 1, ig%n(1)
    ig%at(i)%o%force(1:3) = ig%at(i)%o%force(1:3) + pl%F(:, i)
    ig%at(i)%o%epot = ig%at(i)%o%epot + pl%e(i)
  enddo
  ig%epot = et
end subroutine

! smatb_interact (TB.F90:457-586): the smatb matrices and the global ->
! internal type map of the group (TB.F90:66-102).
subroutine tb_hip_interact(ig, nz, zmap, ca, eps, p, q, r0, rci, rce)
  class(intergroup), intent(inout) :: ig
  integer, intent(in) :: nz, zmap(:)
  real(dp), intent(in) :: ca(nz,nz), eps(nz,nz), p(nz,nz), q(nz,nz), r0(nz,nz), rci(nz,nz), &
                          rce(nz,nz)
  integer(c_int), allocatable :: t(:)
  integer :: i
  if (pack_positions(tb, ig)) then
    allocate(t(ig%n(1)))
    do i = 1, ig%n(1)
      t(i) = zmap(ig%at(i)%o%z) - 1
    enddo
    call new_context(tb, ig, t)
    call check(tb, gh_set_smatb(tb%ctx, int(nz, c_int), ca, eps, p, q, r0, rci, rce))
  else
    call check(tb, gh_set_positions(tb%ctx, tb%x))
  endif
  call compute(tb, ig)
end subroutine

! lj_interact (Pairs.F90:207-279): epsilon (eV), sigma and the cut radius
! (A) of lj_set (Pairs.F90:198-205).
subroutine lj_hip_interact(ig, eps, sigma, rc)
  class(intergroup), intent(inout) :: ig
  real(dp), intent(in) :: eps, sigma, rc
  integer(c_int), allocatable :: t(:)
  if (pack_positions(lj, ig)) then
    allocate(t(ig%n(1)))
    t = 0
    call new_context(lj, ig, t)
    call check(lj, gh_set_lj(lj%ctx, real(eps, c_double), real(sigma, c_double), &
                             real(rc, c_double)))
  else
    call check(lj, gh_set_positions(lj%ctx, lj%x))
  endif
  call compute(lj, ig)
end subroutine

end module gems_plugin_hip
