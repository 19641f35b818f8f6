! gems_hip.f90 -- ISO_C_BINDING interface to the MI355X force-evaluation
! library (include/gems_hip.h).  A GEMS build uses it from Fortran exactly as
! the C header describes: plain values and contiguous real(c_double) /
! integer(c_int) arrays in GEMS internal units (Angstrom, amu, ps, ui).
!
! Entry points and the GEMS routines they replace:
!   gh_compute          interact()            src/Interaction.f90:135-241
!   gh_set_positions    ig%at(i)%o%pos        src/Neighbour.F90:114
!   gh_set_accelerations atom%acel            src/Atoms.F90:70
!   gh_set_smatb        smatb_set/readprm     src/TB.F90:106-170, 291-365
!   gh_set_lj           lj_set                src/Pairs.F90:198-205
!   gh_set_plj          plj_set               src/Pairs.F90:282-291
!   gh_set_slj          slj_set               src/Pairs.F90:374-401
!   gh_set_wca          wca_set               src/Pairs.F90:495-506
!   gh_set_skin         nb_dcut               src/Neighbour.F90:237
!   gh_set_box          box_setvars/set_pbc   src/Program_Types.F90:126-141
!   gh_neighbors        ig%nn, ig%list        src/Neighbour.F90:120-121
!   gh_integrator_vv    evolve v_verlet       src/Integration.F90:1150-1186
!   gh_integrator_ermak evolve ermak T g      src/Integration.F90:909-1031
!   gh_run              dinamica n            src/Programs.f90:114-167
!   gh_prng_lcg         prng lcg / seed       src/Random.F90:137-172
!   gh_get/set_lcg_state lcg chp record       src/Random.F90:178-190
!   gh_lbfgs            lbfgs                 src/Quasi_Newton.f90:160-232
!   gh_set_gdist        set tempgdist T       src/Set_Properties.f90:301-377
module gems_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: GH_FP64 = 0, GH_MIXED = 1
  integer(c_int), parameter :: GH_RNG_LCG = 0, GH_RNG_PHILOX = 1
  integer(c_int), parameter :: GH_POT_SMATB = 1, GH_POT_LJ = 2, GH_POT_PLJ = 3, GH_POT_SLJ = 4, &
                               GH_POT_WCA = 5, GH_POT_SM1 = 6

  ! gh_stats of the C header
  type, bind(C) :: gh_stats
    integer(c_int64_t) :: nupd
    integer(c_int32_t) :: maxnn, cap, ncells(3), list_mode
    integer(c_int64_t) :: steps
    integer(c_int64_t) :: launches
    integer(c_int32_t) :: list_build, chunks
  end type

  interface
    integer(c_int) function gh_create(ctx, device, precision) bind(C, name='gh_create')
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device, precision
    end function

    subroutine gh_destroy(ctx) bind(C, name='gh_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    type(c_ptr) function gh_last_error_c(ctx) bind(C, name='gh_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function gh_set_precision(ctx, precision) bind(C, name='gh_set_precision')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
    end function

    integer(c_int) function gh_set_box(ctx, L, pbc) bind(C, name='gh_set_box')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: L(3)
      integer(c_int), intent(in) :: pbc(3)
    end function

    integer(c_int) function gh_set_box_free(ctx, L, pbc) bind(C, name='gh_set_box')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, L, pbc     ! pass c_null_ptr: no box
    end function

    integer(c_int) function gh_set_sm1(ctx, r1, r2, e) bind(C, name='gh_set_sm1')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: r1, r2, e
    end function

    integer(c_int) function gh_set_images(ctx, ghost) bind(C, name='gh_set_images')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: ghost          ! 0: box mic T, 1: box mic F
    end function

    integer(c_int) function gh_set_pair_masks(ctx, feel, felt) bind(C, name='gh_set_pair_masks')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: feel(*), felt(*)   ! per atom: i pairs j if iand(feel(i), felt(j)) /= 0
    end function

    integer(c_int) function gh_set_atoms(ctx, n, x, type) bind(C, name='gh_set_atoms')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      real(c_double), intent(in) :: x(3, n)
      integer(c_int), intent(in) :: type(n)
    end function

    integer(c_int) function gh_set_positions(ctx, x) bind(C, name='gh_set_positions')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: x(*)
    end function

    integer(c_int) function gh_set_masses(ctx, nt, mass) bind(C, name='gh_set_masses')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt
      real(c_double), intent(in) :: mass(nt)
    end function

    integer(c_int) function gh_set_velocities(ctx, v) bind(C, name='gh_set_velocities')
      import :: c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: v(*)
    end function

    integer(c_int) function gh_set_accelerations(ctx, a) bind(C, name='gh_set_accelerations')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: a(*)
    end function

    integer(c_int) function gh_set_smatb(ctx, nt, A, xi, p, q, r0, rci, rce) &
        bind(C, name='gh_set_smatb')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt
      real(c_double), intent(in) :: A(nt, nt), xi(nt, nt), p(nt, nt), q(nt, nt), &
                                    r0(nt, nt), rci(nt, nt), rce(nt, nt)
    end function

    integer(c_int) function gh_set_lj(ctx, eps, sigma, rc) bind(C, name='gh_set_lj')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      real(c_double), value :: eps, sigma, rc
    end function

    integer(c_int) function gh_set_plj(ctx, nt, eps, sigma, rc) bind(C, name='gh_set_plj')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt
      real(c_double), intent(in) :: eps(nt, nt), sigma(nt, nt)
      real(c_double), value :: rc
    end function

    integer(c_int) function gh_set_slj(ctx, nt, eps, sigma, rs) bind(C, name='gh_set_slj')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt
      real(c_double), intent(in) :: eps(nt, nt), sigma(nt, nt), rs(nt, nt)
    end function

    integer(c_int) function gh_set_wca(ctx, nt, eps, sigma) bind(C, name='gh_set_wca')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt
      real(c_double), intent(in) :: eps(nt, nt), sigma(nt, nt)
    end function

    integer(c_int) function gh_set_skin(ctx, skin) bind(C, name='gh_set_skin')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      real(c_double), value :: skin
    end function

    ! cell-tile shape of the device passes (no reference counterpart)
    integer(c_int) function gh_set_tile(ctx, bd) bind(C, name='gh_set_tile')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: bd(3)
    end function

    integer(c_int) function gh_autotune(ctx, reps, bd_out, us_out) bind(C, name='gh_autotune')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: reps
      integer(c_int), intent(out) :: bd_out(3)
      real(c_double), intent(out) :: us_out
    end function

    integer(c_int) function gh_compute(ctx, F, e_atom, e_tot) bind(C, name='gh_compute')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: F(*), e_atom(*), e_tot
    end function

    integer(c_int) function gh_neighbors(ctx, nn, idx, cap) bind(C, name='gh_neighbors')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: nn(*), idx(*)
      integer(c_int), value :: cap
    end function

    integer(c_int) function gh_prng_lcg(ctx, seed) bind(C, name='gh_prng_lcg')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: seed
    end function

    integer(c_int) function gh_get_lcg_state(ctx, state) bind(C, name='gh_get_lcg_state')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: state(3)
    end function

    integer(c_int) function gh_set_lcg_state(ctx, state) bind(C, name='gh_set_lcg_state')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: state(3)
    end function

    integer(c_int) function gh_lbfgs(ctx, m, eps, max_iter, iters, epot_ev) bind(C, name='gh_lbfgs')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: m, max_iter
      real(c_double), value :: eps
      integer(c_int), intent(out) :: iters
      real(c_double), intent(out) :: epot_ev
    end function

    integer(c_int) function gh_set_gdist(ctx, T) bind(C, name='gh_set_gdist')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      real(c_double), value :: T
    end function

    integer(c_int) function gh_integrator_vv(ctx, dt) bind(C, name='gh_integrator_vv')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function

    integer(c_int) function gh_integrator_ermak(ctx, dt, T, gamma, rng, seed) &
        bind(C, name='gh_integrator_ermak')
      import :: c_ptr, c_double, c_int, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), value :: dt, T, gamma
      integer(c_int), value :: rng
      integer(c_int64_t), value :: seed
    end function

    integer(c_int) function gh_run(ctx, nsteps, each, energy) bind(C, name='gh_run')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps, each
      real(c_double), intent(out) :: energy(*)
    end function

    integer(c_int) function gh_energies(ctx, epot, ekin) bind(C, name='gh_energies')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: epot, ekin
    end function

    integer(c_int) function gh_get_state(ctx, x, v, F, e) bind(C, name='gh_get_state')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: x(*), v(*), F(*), e(*)
    end function
    ! ---- batched free clusters (gh_batch_*): replica-exchange ladders and
    !      FIRE relaxations of many systems in one launch
    integer(c_int) function gh_batch_create(b, device, precision) bind(C, name='gh_batch_create')
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: b
      integer(c_int), value :: device, precision
    end function

    subroutine gh_batch_destroy(b) bind(C, name='gh_batch_destroy')
      import :: c_ptr
      type(c_ptr), value :: b
    end subroutine

    integer(c_int) function gh_batch_set_systems(b, nsys, off, x, type) &
        bind(C, name='gh_batch_set_systems')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: nsys
      integer(c_int), intent(in) :: off(nsys + 1)
      real(c_double), intent(in) :: x(*)
      integer(c_int), intent(in) :: type(*)
    end function

    integer(c_int) function gh_batch_set_smatb(b, nt, A, xi, p, q, r0, rci, rce) &
        bind(C, name='gh_batch_set_smatb')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: nt
      real(c_double), intent(in) :: A(nt, nt), xi(nt, nt), p(nt, nt), q(nt, nt), &
                                    r0(nt, nt), rci(nt, nt), rce(nt, nt)
    end function

    integer(c_int) function gh_batch_set_masses(b, nt, mass) bind(C, name='gh_batch_set_masses')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: nt
      real(c_double), intent(in) :: mass(nt)
    end function

    integer(c_int) function gh_batch_set_temperatures(b, T) &
        bind(C, name='gh_batch_set_temperatures')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: b
      real(c_double), intent(in) :: T(*)
    end function

    integer(c_int) function gh_batch_langevin(b, nsteps, dt, gamma, seed, epot, ekin) &
        bind(C, name='gh_batch_langevin')
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: b
      integer(c_int), value :: nsteps
      real(c_double), value :: dt, gamma
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: epot(*), ekin(*)
    end function

    integer(c_int) function gh_batch_exchange(b, nladder, R, sys, parity, seed, naccept) &
        bind(C, name='gh_batch_exchange')
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: b
      integer(c_int), value :: nladder, R, parity
      integer(c_int), intent(inout) :: sys(*)
      integer(c_int64_t), value :: seed
      integer(c_int), intent(out) :: naccept
    end function

    integer(c_int) function gh_batch_fire(b, max_iter, dt_max, alpha0, fswitch, ftol, iters, &
                                          converged, epot, fmax) bind(C, name='gh_batch_fire')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: max_iter
      real(c_double), value :: dt_max, alpha0, fswitch, ftol
      integer(c_int), intent(out) :: iters(*), converged(*)
      real(c_double), intent(out) :: epot(*), fmax(*)
    end function

    integer(c_int) function gh_batch_get_state(b, x, v) bind(C, name='gh_batch_get_state')
      import :: c_ptr, c_double, c_int
      type(c_ptr), value :: b
      real(c_double), intent(out) :: x(*), v(*)
    end function

    ! device arrays are passed as c_ptr (their device addresses); no allocation
    integer(c_int) function gh_cluster_forces_static(device, precision, n, nt, x, type, a, xi, &
        p, q, r0, rci, rce, skin, f, e_atom, epot, status, stream) &
        bind(C, name='gh_cluster_forces_static')
      import :: c_ptr, c_double, c_int
      integer(c_int), value :: device, precision, n, nt
      type(c_ptr), value :: x, type
      real(c_double), intent(in) :: a(*), xi(*), p(*), q(*), r0(*), rci(*), rce(*)
      real(c_double), value :: skin
      type(c_ptr), value :: f, e_atom, epot, status, stream
    end function
    integer(c_int) function gh_get_stats(ctx, st) bind(C, name='gh_get_stats')
      import :: c_ptr, c_int, gh_stats
      type(c_ptr), value :: ctx
      type(gh_stats), intent(out) :: st
    end function

    integer(c_int) function gh_set_timing(ctx, enable) bind(C, name='gh_set_timing')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function

    integer(c_int) function gh_get_timing(ctx, ms, count) bind(C, name='gh_get_timing')
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(4)
      integer(c_int64_t), intent(out) :: count(4)
    end function

    type(c_ptr) function gh_batch_last_error_c(b) bind(C, name='gh_batch_last_error')
      import :: c_ptr
      type(c_ptr), value :: b
    end function

    integer(c_int) function gh_batch_set_precision(b, precision) bind(C, name='gh_batch_set_precision')
      import :: c_ptr, c_int
      type(c_ptr), value :: b
      integer(c_int), value :: precision
    end function

    integer(c_int) function gh_batch_set_skin(b, skin) bind(C, name='gh_batch_set_skin')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: b
      real(c_double), value :: skin
    end function

    integer(c_int) function gh_batch_set_velocities(b, v) bind(C, name='gh_batch_set_velocities')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: b
      real(c_double), intent(in) :: v(*)        ! 3 x (all atoms of all systems)
    end function

    integer(c_int) function gh_batch_compute(b, F, epot) bind(C, name='gh_batch_compute')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: b
      real(c_double), intent(out) :: F(*)       ! 3 x (all atoms), ui/A
      real(c_double), intent(out) :: epot(*)    ! per system, eV
    end function
  end interface

contains

  ! gh_last_error as a Fortran string (for werr).
  function gh_last_error(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    interface
      integer(c_size_t) function c_strlen(str) bind(C, name='strlen')
        import :: c_ptr, c_size_t
        type(c_ptr), value :: str
      end function
    end interface
    p = gh_last_error_c(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

end module gems_hip
