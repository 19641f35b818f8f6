! gems_plugin_hip.f90 -- the force calls of a GEMS build, on the MI355X
! library.  GEMS keeps its own script interpreter, outputs and atom lists;
! when the environment selects the device backend (GEMS_BACKEND=hip),
! smatb_interact (src/TB.F90:457-586) hands each force evaluation to
! tb_hip_interact, and lj_interact, plj_interact, slj_interact and
! wca_interact (src/Pairs.F90:207-560) to pair_hip_interact.  Each of them
!   * records the group's potential (its kind and parameters) under the
!     group's igr_vop index (ig%id, Neighbour.F90:268-272), for the device
!     MD loop (gems_plugin_hip_md) to build its context from,
!   * serves self (`under`) groups when no virial is asked for
!     (b_gvirial/b_avirial, Program_Types.F90:78); otherwise it returns
!     .false. and GEMS computes the group on the host as before,
!   * keeps one device context per served group, rebuilt when the group's
!     atoms, types, the potential's parameters, the box (box,
!     Program_Types.F90:69) or the periodic-image mode change,
!   * packs the group's atom positions ig%at(i)%o%pos (Neighbour.F90:114)
!     and afterwards only sends positions (gh_set_positions: the device
!     keeps its lists and decides rebuilds itself), so the group's own list
!     builder is switched off (lista => intergroup0_empty, Neighbour.F90:
!     669-672; GEMS then builds no list for it),
!   * calls gh_compute and accumulates forces (ui/A) and per-atom energies
!     (ui) into the atoms as the reference routines do, and sets ig%epot (ui).
! GEMS_HIP_PRECISION=mixed selects the mixed-precision form (default FP64).
! Both periodic-image modes are served (mic, Program_Types.F90:98): `box mic
! T` pairs nearest images, `box mic F` every first-shell image as GEMS's ghost
! atoms do (gh_set_images); the device builds its own images either way and
! only the group's local atoms (ig%n(1)) travel.
!
! hip_system_ctx builds the context of the whole interaction set for the
! device MD loop: every enabled group's recorded potential on the integrated
! atoms, in the integrator's atom order.  SMA-TB and lj: one self group over
! every atom.  plj, slj, wca: any number of self and `with` groups of one
! kind; each atom's type is the set of group lists it belongs to (atom%igr,
! atom%igra, Atoms.F90:105-107), and each pair of types takes the parameters
! of the one group that pairs them (plj and wca: no group, no interaction).
module gems_plugin_hip
use iso_c_binding
use gems_hip
use gems_constants, only: dp, dm
use gems_program_types, only: box, mic, b_gvirial, b_avirial
use gems_atoms, only: atom_ap
use gems_neighbour, only: intergroup, nb_dcut, intergroup0_empty, igr_vop
use gems_errors, only: werr
implicit none
private
public :: gems_hip_on, tb_hip_interact, pair_hip_interact, hip_system_ctx
public :: HIP_TB, HIP_LJ, HIP_PLJ, HIP_SLJ, HIP_WCA

integer, parameter :: HIP_TB = 1, HIP_LJ = 2, HIP_PLJ = 3, HIP_SLJ = 4, HIP_WCA = 5
integer, parameter :: MAXPL = 64, MAXT = 8

! the potential of one interaction group, as its force routine last saw it
type recipe
  integer :: kind = 0
  integer :: nz = 0
  integer, allocatable :: zmap(:)           ! TB: global element -> internal type
  real(dp), allocatable :: par(:)           ! TB: ca eps p q r0 rci rce (nz x nz)
                                            ! pairs: e s rs rcut
end type

! device state of one interaction group (or of the whole set: sys)
type plugin
  type(c_ptr) :: ctx = c_null_ptr
  integer :: nat = -1
  logical :: list_off = .false.             ! GEMS's list builder switched off
  real(dp) :: box_set(dm) = -1.0_dp
  logical :: mic_set = .true.
  real(c_double), allocatable :: par(:)     ! the potential's parameters as set
  integer(c_int), allocatable :: t(:)       ! the atoms' types as set
  type(atom_ap), allocatable :: at(:)       ! the atoms served, in device order
  real(c_double), allocatable :: x(:,:), F(:,:), e(:)
end type

type(recipe), save :: rec(MAXPL)
type(plugin), save, target :: pls(MAXPL), sys
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

! The slot of group ig: its igr_vop index (set by intergroup_constructor).
integer function slot_of(ig)
  class(intergroup), intent(in) :: ig
  call werr('gems_hip: more than 64 interaction groups', ig%id < 1 .or. ig%id > MAXPL)
  slot_of = ig%id
end function

! True when group ig is computed on the device; switches its list off then.
logical function served(ig, k)
  class(intergroup), intent(inout) :: ig
  integer, intent(in) :: k
  served = ig%self .and. .not. (b_gvirial .or. b_avirial)
  if (.not. served) then
    call werr('gems_hip: the virial was asked for after the device took over a group', &
              pls(k)%list_off)
    return
  endif
  ig%lista => intergroup0_empty
  pls(k)%list_off = .true.
end function

! Positions of pl%at in; true when the context must be (re)built (atoms,
! types, parameters, box or image mode changed).
logical function pack_positions(pl, t, par)
  type(plugin), intent(inout) :: pl
  integer(c_int), intent(in) :: t(:)
  real(c_double), intent(in) :: par(:)
  integer :: i, n
  logical :: same
  n = size(pl%at)
  if (allocated(pl%x)) then
    if (size(pl%x, 2) /= n) deallocate(pl%x, pl%F, pl%e)
  endif
  if (.not. allocated(pl%x)) allocate(pl%x(3, n), pl%F(3, n), pl%e(n))
  do i = 1, n
    pl%x(:, i) = pl%at(i)%o%pos(1:3)
  enddo
  same = c_associated(pl%ctx) .and. n == pl%nat .and. all(box(1:3) == pl%box_set(1:3)) .and. &
         (mic .eqv. pl%mic_set)
  if (same) same = size(pl%t) == size(t) .and. size(pl%par) == size(par)
  if (same) same = all(pl%t == t) .and. all(pl%par == par)
  pack_positions = .not. same
end function

! A new context with the box and the atoms pl%at (types t, 0-based).
subroutine new_context(pl, t, par)
  type(plugin), intent(inout) :: pl
  integer(c_int), intent(in) :: t(:)
  real(c_double), intent(in) :: par(:)
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
    if (pl%at(1)%o%pbc(i)) pbc(i) = 1
  enddo
  lbox = box(1:3)
  if (any(pbc /= 0)) then
    call check(pl, gh_set_box(pl%ctx, lbox, pbc))
  else
    call check(pl, gh_set_box_free(pl%ctx, c_null_ptr, c_null_ptr))
  endif
  call check(pl, gh_set_images(pl%ctx, merge(0_c_int, 1_c_int, mic)))
  call check(pl, gh_set_atoms(pl%ctx, int(size(pl%at), c_int), pl%x, t))
  call check(pl, gh_set_skin(pl%ctx, real(nb_dcut, c_double)))
  pl%nat = size(pl%at)
  pl%box_set(1:3) = box(1:3)
  pl%mic_set = mic
  if (allocated(pl%t)) deallocate(pl%t)
  if (allocated(pl%par)) deallocate(pl%par)
  allocate(pl%t(size(t)), pl%par(size(par)))
  pl%t = t
  pl%par = par
end subroutine

! Forces and energies out, accumulated as the reference routines do.
subroutine compute(pl, ig)
  type(plugin), intent(inout) :: pl
  class(intergroup), intent(inout) :: ig
  integer :: i
  real(c_double) :: et
  call check(pl, gh_compute(pl%ctx, pl%F, pl%e, et))
  do i = 1, size(pl%at)
    pl%at(i)%o%force(1:3) = pl%at(i)%o%force(1:3) + pl%F(:, i)
    pl%at(i)%o%epot = pl%at(i)%o%epot + pl%e(i)
  enddo
  ig%epot = et
end subroutine

! The group's local atoms, in its own order.
subroutine take_group_atoms(pl, ig)
  type(plugin), intent(inout) :: pl
  class(intergroup), intent(in) :: ig
  integer :: i
  if (allocated(pl%at)) then
    if (size(pl%at) /= ig%n(1)) deallocate(pl%at)
  endif
  if (.not. allocated(pl%at)) allocate(pl%at(ig%n(1)))
  do i = 1, ig%n(1)
    pl%at(i)%o => ig%at(i)%o
  enddo
end subroutine

subroutine set_smatb(pl, nz, par)
  type(plugin), intent(inout) :: pl
  integer, intent(in) :: nz
  real(dp), intent(in) :: par(:)
  integer :: m
  m = nz * nz
  call check(pl, gh_set_smatb(pl%ctx, int(nz, c_int), par(1:m), par(m+1:2*m), par(2*m+1:3*m), &
             par(3*m+1:4*m), par(4*m+1:5*m), par(5*m+1:6*m), par(6*m+1:7*m)))
end subroutine

! smatb_interact (TB.F90:457-586): the smatb matrices and the global ->
! internal type map of the group (TB.F90:66-102).  .false.: not served.
logical function tb_hip_interact(ig, nz, zmap, ca, eps, p, q, r0, rci, rce) result(ok)
  class(intergroup), intent(inout) :: ig
  integer, intent(in) :: nz, zmap(:)
  real(dp), intent(in) :: ca(nz,nz), eps(nz,nz), p(nz,nz), q(nz,nz), r0(nz,nz), rci(nz,nz), &
                          rce(nz,nz)
  integer(c_int), allocatable :: t(:)
  integer :: i, k
  k = slot_of(ig)
  rec(k)%kind = HIP_TB
  rec(k)%nz = nz
  rec(k)%zmap = zmap
  rec(k)%par = [reshape(ca, [nz*nz]), reshape(eps, [nz*nz]), reshape(p, [nz*nz]), &
                reshape(q, [nz*nz]), reshape(r0, [nz*nz]), reshape(rci, [nz*nz]), &
                reshape(rce, [nz*nz])]
  ok = served(ig, k)
  if (.not. ok) return
  call take_group_atoms(pls(k), ig)
  allocate(t(ig%n(1)))
  do i = 1, ig%n(1)
    t(i) = zmap(ig%at(i)%o%z) - 1
  enddo
  if (pack_positions(pls(k), t, rec(k)%par)) then
    call new_context(pls(k), t, rec(k)%par)
    call set_smatb(pls(k), nz, rec(k)%par)
  else
    call check(pls(k), gh_set_positions(pls(k)%ctx, pls(k)%x))
  endif
  call compute(pls(k), ig)
end function

subroutine set_pairs(pl, kind, nt, e, s, rs, rc)
  type(plugin), intent(inout) :: pl
  integer, intent(in) :: kind, nt
  real(c_double), intent(in) :: e(:), s(:), rs(:), rc
  select case (kind)
  case (HIP_LJ)
    call check(pl, gh_set_lj(pl%ctx, e(1), s(1), rc))
  case (HIP_PLJ)
    call check(pl, gh_set_plj(pl%ctx, int(nt, c_int), e, s, rc))
  case (HIP_SLJ)
    call check(pl, gh_set_slj(pl%ctx, int(nt, c_int), e, s, rs))
  case (HIP_WCA)
    call check(pl, gh_set_wca(pl%ctx, int(nt, c_int), e, s))
  end select
end subroutine

! lj_interact, plj_interact, slj_interact, wca_interact (Pairs.F90:207-560):
! epsilon (eV), sigma (A), slj's spline point rs (A; 0 for the others) and
! the group's cut radius (ig%rcut: lj and plj use it).  .false.: not served.
logical function pair_hip_interact(ig, kind, e, s, rs) result(ok)
  class(intergroup), intent(inout) :: ig
  integer, intent(in) :: kind
  real(dp), intent(in) :: e, s, rs
  integer(c_int), allocatable :: t(:)
  integer :: k
  k = slot_of(ig)
  rec(k)%kind = kind
  rec(k)%par = [e, s, rs, ig%rcut]
  ok = served(ig, k)
  if (.not. ok) return
  call take_group_atoms(pls(k), ig)
  allocate(t(ig%n(1)))
  t = 0
  if (pack_positions(pls(k), t, rec(k)%par)) then
    call new_context(pls(k), t, rec(k)%par)
    call set_pairs(pls(k), kind, 1, [e], [s], [rs], real(ig%rcut, c_double))
  else
    call check(pls(k), gh_set_positions(pls(k)%ctx, pls(k)%x))
  endif
  call compute(pls(k), ig)
end function

! The device context of every enabled interaction group acting on the atoms
! ap (the integrator's atoms, in its order), created or brought up to date
! (positions sent).  Null, with the reason in why, when the set is not one
! the device serves.  Per-type masses are set from the atoms; the atoms of
! a type must share one mass.
function hip_system_ctx(ap, why) result(ctx)
  type(atom_ap), intent(in) :: ap(:)
  character(len=:), allocatable, intent(out) :: why
  type(c_ptr) :: ctx
  class(intergroup), pointer :: ig
  integer :: i, j, g, n, kind, ng, nt, s1, s2, gid(MAXPL)
  integer(c_int), allocatable :: t(:)
  integer(8) :: code, codes(MAXT)
  real(c_double), allocatable :: par(:)
  real(c_double) :: e(MAXT*MAXT), s(MAXT*MAXT), rs(MAXT*MAXT), rc, mass(MAXT)
  logical :: set(MAXT*MAXT), ina(MAXT), inb(MAXT), same
  ctx = c_null_ptr
  n = size(ap)
  why = ''
  ! the enabled groups and their common kind
  ng = 0
  kind = 0
  do i = 1, igr_vop%size
    ig => igr_vop%o(i)%o
    if (ig%disable) cycle
    if (ig%id < 1 .or. ig%id > MAXPL) then
      why = 'more than 64 interaction groups'
      return
    endif
    if (rec(ig%id)%kind == 0) then
      why = 'an interaction group of a kind the device does not compute'
      return
    endif
    if (kind /= 0 .and. rec(ig%id)%kind /= kind) then
      why = 'interaction groups of more than one kind'
      return
    endif
    kind = rec(ig%id)%kind
    if (.not. ig%self .and. .not. ig%newton) then
      why = 'a one-sided (`feels`) interaction group'
      return
    endif
    ng = ng + 1
    gid(ng) = ig%id
  enddo
  if (ng == 0) then
    why = 'no interaction group'
    return
  endif
  if (b_gvirial .or. b_avirial) then
    why = 'the virial is asked for'
    return
  endif
  allocate(t(n))
  nt = 1
  if (kind == HIP_TB .or. kind == HIP_LJ) then
    ig => igr_vop%o(gid(1))%o
    if (ng /= 1 .or. .not. ig%self .or. ig%n(1) /= n) then
      why = 'SMA-TB and lj are served as one self group over every integrated atom'
      return
    endif
    t = 0
    if (kind == HIP_TB) then
      do i = 1, n
        t(i) = rec(gid(1))%zmap(ap(i)%o%z) - 1
      enddo
      nt = rec(gid(1))%nz
    else
      e(1) = rec(gid(1))%par(1)
      s(1) = rec(gid(1))%par(2)
      rs(1) = 0.0_dp
      rc = rec(gid(1))%par(4)
    endif
    par = [rec(gid(1))%par, real(kind, dp)]
  else
    ! types: the set of group lists each atom belongs to
    if (ng > 31) then
      why = 'more than 31 pair groups'
      return
    endif
    nt = 0
    do i = 1, n
      code = 0
      do j = 1, ap(i)%o%nigr
        do g = 1, ng
          if (ap(i)%o%igr(j) /= gid(g)) cycle
          if (ap(i)%o%igra(j)) then
            code = ibset(code, 2 * (g - 1))
          else
            code = ibset(code, 2 * (g - 1) + 1)
          endif
        enddo
      enddo
      t(i) = -1
      do j = 1, nt
        if (codes(j) == code) t(i) = j - 1
      enddo
      if (t(i) < 0) then
        if (nt == MAXT) then
          why = 'more than 8 atom types across the pair groups'
          return
        endif
        nt = nt + 1
        codes(nt) = code
        t(i) = nt - 1
      endif
    enddo
    ! each pair of types takes the parameters of the one group pairing them
    set = .false.
    e = 0.0_dp
    s = 1.0_dp
    rs = 1.0_dp
    rc = rec(gid(1))%par(4)
    do g = 1, ng
      ig => igr_vop%o(gid(g))%o
      do j = 1, nt
        ina(j) = btest(codes(j), 2 * (g - 1))
        inb(j) = btest(codes(j), 2 * (g - 1) + 1)
      enddo
      if (.not. ig%self .and. any(ina(1:nt) .and. inb(1:nt))) then
        why = 'an atom on both sides of a `with` group'
        return
      endif
      if (kind == HIP_PLJ .and. rec(gid(g))%par(4) /= rc) then
        why = 'plj groups with different cut radii'
        return
      endif
      do s1 = 1, nt
        do s2 = 1, nt
          if (ig%self) then
            if (.not. (ina(s1) .and. ina(s2))) cycle
          else
            if (.not. ((ina(s1) .and. inb(s2)) .or. (inb(s1) .and. ina(s2)))) cycle
          endif
          j = (s1 - 1) * nt + s2
          if (set(j)) then
            why = 'two interaction groups act on the same pair of atoms'
            return
          endif
          set(j) = .true.
          e(j) = rec(gid(g))%par(1)
          s(j) = rec(gid(g))%par(2)
          rs(j) = rec(gid(g))%par(3)
        enddo
      enddo
    enddo
    if (kind == HIP_SLJ .and. .not. all(set(1:nt*nt))) then
      why = 'slj groups that leave a pair of atom types without a potential'
      return
    endif
    par = [e(1:nt*nt), s(1:nt*nt), rs(1:nt*nt), rc, real(kind, dp), real(nt, dp)]
  endif
  ! masses per type
  mass = 1.0_dp
  do i = 1, n
    mass(t(i) + 1) = ap(i)%o%mass
  enddo
  do i = 1, n
    if (ap(i)%o%mass /= mass(t(i) + 1)) then
      why = 'atoms of one type with different masses'
      return
    endif
  enddo
  ! the context: reused while the atoms (by identity), types and potential stay
  same = allocated(sys%at)
  if (same) same = size(sys%at) == n
  if (same) then
    do i = 1, n
      if (.not. associated(sys%at(i)%o, ap(i)%o)) then
        same = .false.
        exit
      endif
    enddo
  endif
  if (.not. same) then
    if (allocated(sys%at)) deallocate(sys%at)
    allocate(sys%at(n))
    do i = 1, n
      sys%at(i)%o => ap(i)%o
    enddo
    sys%nat = -1
  endif
  if (pack_positions(sys, t, par)) then
    call new_context(sys, t, par)
    if (kind == HIP_TB) then
      call set_smatb(sys, nt, rec(gid(1))%par)
    else
      call set_pairs(sys, kind, nt, e(1:nt*nt), s(1:nt*nt), rs(1:nt*nt), rc)
    endif
  else
    call check(sys, gh_set_positions(sys%ctx, sys%x))
  endif
  call check(sys, gh_set_masses(sys%ctx, int(nt, c_int), mass(1:nt)))
  ctx = sys%ctx
end function

end module gems_plugin_hip
