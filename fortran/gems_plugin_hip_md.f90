! gems_plugin_hip_md.f90 -- GEMS's `dinamica` loop (dinamic,
! src/Programs.f90:114-167) on the device.  With GEMS_BACKEND=hip, dinamic
! calls hip_dinamic after its first interact(.false.); when the simulation is
! one the device can step on its own, the steps run there in chunks (gh_run)
! and GEMS sees the state only where the loop would look at it:
!   * after each step whose number is a multiple of an open outfile's
!     sumeach (write_out_real, Output.F90:521-545: nothing is written at
!     other steps), of chpeach when checkpoints are on (b_ckp), and at the
!     powers of ten that timer_dump prints (Errors.f90:128-172);
! there the atoms get positions (wrapped as do_pbc_1 does, Set_Properties.f90:
! 444-470, with boxcr advanced by the nearest-image displacement since the
! last look, which needs |displacement| < L/2 over one chunk: chunks are cut
! at 1000 steps in periodic boxes), velocities, forces, per-atom energies and
! acel = F/m (velocity_verlet_b and ermak_b leave that), tepot and the group
! energies are set, dm_steps and time advance as the loop advances them
! (time + dt once per step), ghost images follow (pbcghost_move), and
! write_out(1, .), write_chp and timer_dump run as in the loop.  GEMS's
! nupd_vlist counts the device's list rebuilds.
! Served: one integrator (its, Integration.F90:63) over every local atom,
! `evolve v_verlet` (velocity_verlet_a/b, Integration.F90:1150-1186) or
! `evolve ermak T gamma` (ermak_a/b, :963-1031) with `prng lcg` (its stream
! continues on the device: Random.F90:145-161, state handed over both ways);
! every enabled interaction group one the device computes (hip_system_ctx);
! no force field bonds (boundgrs), collective variables (ncvs), bias,
! constrained atoms (bconst), command interpreter blocks (b_load) or outputs
! written inside the force call (enable(2..)).  Otherwise, or with
! GEMS_HIP_LOOP=0, hip_dinamic returns .false. and GEMS runs its own loop
! (the force calls still go to the device where gems_plugin_hip serves them);
! the log says which.  With several pair groups the device knows only the
! total energy: each group's epot (the `eparts` columns) is left at zero.
module gems_plugin_hip_md
use iso_c_binding
use gems_hip
use gems_plugin_hip, only: gems_hip_on, hip_system_ctx
use gems_constants, only: dp, dm
use gems_program_types, only: dm_steps, time, dt, box, nlocal, ghost
use gems_atoms, only: atom, atom_ap, atom_dclist
use gems_neighbour, only: intergroup, igr_vop, pbcghost_move, nupd_vlist
use gems_integration, only: its, its_hip_kind
use gems_random, only: lcg_hip_state
use gems_interaction, only: tepot
use gems_bias, only: biason, cepot, bias
use gems_cvs, only: ncvs
use gems_forcefield, only: boundgrs
use gems_input_parsing, only: b_load
use gems_output, only: write_out, of_vop, outfile
use gems_checkpoint, only: b_ckp, chpeach, write_chp
use gems_set_properties, only: posvel_changed, epot_changed
use gems_errors, only: werr, wlog, timer_dump
implicit none
private
public :: hip_dinamic

contains

logical function hip_dinamic(steps, b_out, b_time) result(done)
  integer, intent(in) :: steps
  logical, intent(in) :: b_out, b_time
  character(len=:), allocatable :: why
  character(len=32) :: v
  character(len=160) :: msg
  type(atom_ap), allocatable :: ap(:)
  type(atom_dclist), pointer :: la
  class(atom), pointer :: o
  class(intergroup), pointer :: ig
  type(outfile), pointer :: of
  type(c_ptr) :: ctx
  type(gh_stats) :: st
  real(c_double), allocatable :: x(:,:), vel(:,:), F(:,:), e(:)
  real(c_double) :: dummy(3)
  integer(c_int) :: rc, lcg(3)
  integer(8) :: m0, m, s, nxt, se, p10, k, kdone, s0, nupd0
  integer :: i, j, kind, nat, l, ist, tries, ngr
  integer(8) :: c0, c1, crate, ca, cb, tdev, tcopy, tback, tgems
  real(dp) :: ibox(3)
  logical :: periodic, ok

  done = .false.
  if (.not. gems_hip_on()) return
  call get_environment_variable('GEMS_HIP_LOOP', v, l, ist)
  if (ist == 0 .and. l >= 1) then
    if (v(1:1) == '0') return
  endif

  ! ---- what the loop would do besides integrating and computing forces
  why = ''
  kind = 0
  if (its%size /= 1) then
    why = 'not exactly one integrator'
  else
    kind = its_hip_kind(1)
    nat = its%o(1)%nat
    if (kind == 0) then
      why = 'an integrator other than v_verlet and ermak'
    else if (nat /= nlocal .or. nat < 1) then
      why = 'the integrator does not move every atom'
    else if (kind == 2 .and. its%o(1)%p%o(9) /= dt) then
      why = 'ermak set with another time step'
    endif
  endif
  if (why == '' .and. associated(boundgrs%next)) why = 'force field bonds'
  if (why == '' .and. ncvs > 0) why = 'collective variables'
  if (why == '' .and. biason) why = 'a bias'
  if (why == '' .and. b_load) why = 'a command block run during the loop'
  if (why == '' .and. b_out) then
    do i = 1, of_vop%size
      of => of_vop%o(i)%o
      if (.not. of%open) cycle
      if (any(of%enable(2:))) why = 'an output written inside the force call'
    enddo
  endif
  if (why == '') then
    allocate(ap(nat))
    la => its%o(1)%alist
    do i = 1, nat
      la => la%next
      ap(i)%o => la%o
      if (la%o%bconst) why = 'constrained atoms'
    enddo
  endif
  if (why == '' .and. kind == 2) then
    call lcg_hip_state(lcg, .false., ok)
    if (.not. ok) why = 'ermak without `prng lcg`'
  endif
  if (why == '') then
    ctx = hip_system_ctx(ap, why)
  endif
  if (why /= '') then
    call wlog('HIP', 'dinamica on the host: '//why)
    return
  endif

  ! ---- hand the state over
  allocate(x(3, nat), vel(3, nat), F(3, nat), e(nat))
  do i = 1, nat
    vel(:, i) = ap(i)%o%vel(1:3)
    F(:, i) = ap(i)%o%acel(1:3)
  enddo
  call check(gh_set_velocities(ctx, vel))
  call check(gh_set_accelerations(ctx, F))
  if (kind == 1) then
    call check(gh_integrator_vv(ctx, real(dt, c_double)))
  else
    call check(gh_integrator_ermak(ctx, real(dt, c_double), real(its%o(1)%fixt, c_double), &
                                   real(its%o(1)%p%o(10), c_double), GH_RNG_LCG, 0_c_int64_t))
    call check(gh_set_lcg_state(ctx, lcg))
  endif
  periodic = any(ap(1)%o%pbc(1:3))
  ngr = 0
  do j = 1, igr_vop%size
    if (.not. igr_vop%o(j)%o%disable) ngr = ngr + 1
  enddo
  call check(gh_get_stats(ctx, st))
  nupd0 = st%nupd
  write(msg, '(a,i0,a,i0,a)') 'dinamica ', steps, ' steps of ', nat, ' atoms on the device'
  call wlog('HIP', trim(msg))

  ! ---- the loop, in chunks between the steps GEMS looks at
  call system_clock(c0, crate)
  tdev = 0
  tcopy = 0
  tback = 0
  tgems = 0
  ibox = 1.0_dp / box(1:3)
  m0 = nint(dm_steps, 8)
  m = m0
  do while (m < m0 + steps)
    nxt = m0 + steps
    if (periodic) nxt = min(nxt, m + 1000)
    if (b_out) then
      do i = 1, of_vop%size
        of => of_vop%o(i)%o
        if (.not. of%open .or. .not. of%enable(1)) cycle
        se = max(1, of%sumeach)
        nxt = min(nxt, (m / se + 1) * se)
      enddo
    endif
    if (b_ckp) then
      se = max(1, chpeach)
      nxt = min(nxt, (m / se + 1) * se)
    endif
    if (b_time) then
      p10 = 1
      do while (m0 + p10 <= m)
        p10 = p10 * 10
      enddo
      nxt = min(nxt, m0 + p10)
    endif
    k = nxt - m
    ! a stop for a list or halo overflow re-sizes; the next call completes the
    ! stopped step and goes on
    call system_clock(ca)
    call check(gh_get_stats(ctx, st))
    s0 = st%steps
    kdone = 0
    tries = 0
    do while (kdone < k)
      rc = gh_run(ctx, int(k - kdone, c_int), 0_c_int, dummy)
      call check(gh_get_stats(ctx, st))
      kdone = st%steps - s0
      if (rc == 0) cycle
      tries = tries + 1
      if (rc /= -4 .or. tries > 20) call check(rc)
    enddo
    call system_clock(cb)
    tdev = tdev + (cb - ca)
    ! ---- the state where the loop looks at it
    call check(gh_get_state(ctx, x, vel, F, e))
    call system_clock(ca)
    tcopy = tcopy + (ca - cb)
    tepot = 0.0_dp
    do i = 1, nat
      o => ap(i)%o
      do j = 1, 3
        if (o%pbc(j)) o%boxcr(j) = o%boxcr(j) + nint((o%pos(j) - x(j, i)) * ibox(j))
        o%pos(j) = x(j, i)
        o%vel(j) = vel(j, i)
        o%force(j) = F(j, i)
        o%acel(j) = F(j, i) * o%one_mass
      enddo
      o%epot = e(i)
      tepot = tepot + e(i)
    enddo
    cepot = tepot
    bias = 0.0_dp
    do j = 1, igr_vop%size
      ig => igr_vop%o(j)%o
      if (ig%disable) cycle
      ig%epot = merge(tepot, 0.0_dp, ngr == 1)
    enddo
    if (kind == 2) then
      call check(gh_get_lcg_state(ctx, lcg))
      call lcg_hip_state(lcg, .true., ok)
    endif
    if (ghost) call pbcghost_move()
    call posvel_changed()
    call epot_changed()
    nupd_vlist = nupd_vlist + int(st%nupd - nupd0)
    nupd0 = st%nupd
    call system_clock(cb)
    tback = tback + (cb - ca)
    ca = cb
    do s = m + 1, nxt
      time = time + dt
    enddo
    m = nxt
    dm_steps = real(m, dp)
    if (b_out) call write_out(1, dm_steps)
    if (b_ckp) then
      if (mod(m, int(max(1, chpeach), 8)) == 0) call write_chp(int(m - m0), steps)
    endif
    if (b_time) call timer_dump(int(m - m0), time, nupd_vlist)
    call system_clock(cb)
    tgems = tgems + (cb - ca)
  enddo
  call system_clock(c1)
  write(msg, '(a,f0.3,a,es10.3,a)') 'dinamica done in ', real(c1 - c0, dp) / real(crate, dp), &
      ' s (', real(steps, dp) * nat * real(crate, dp) / real(max(1_8, c1 - c0), dp), &
      ' atom-steps/s with the state written back where GEMS looks at it)'
  call wlog('HIP', trim(msg))
  write(msg, '(a,f0.3,a,f0.3,a,f0.3,a,f0.3,a)') 'dinamica time: device steps ', &
      real(tdev, dp) / real(crate, dp), ' s, state copies ', real(tcopy, dp) / real(crate, dp), &
      ' s, state into the atoms ', real(tback, dp) / real(crate, dp), ' s, GEMS output ', &
      real(tgems, dp) / real(crate, dp), ' s'
  call wlog('HIP', trim(msg))
  done = .true.

contains

  subroutine check(rc)
    integer(c_int), intent(in) :: rc
    if (rc /= 0) call werr('gems_hip: '//gh_last_error(ctx), .true.)
  end subroutine

end function

end module gems_plugin_hip_md
