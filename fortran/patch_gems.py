"""Writes the hooked copies of GEMS's sources that fortran/gems.mk compiles.

    python3 fortran/patch_gems.py <reference src dir> <output dir>

The reference sources are read in place and never changed; each hooked file
is written to the output directory (a temporary directory of the build) with
these insertions, each of which must apply exactly once:

  Pairs.F90      `use gems_plugin_hip` beside `use gems_neighbour`; at the top
                 of lj_interact, plj_interact, slj_interact, wca_interact
                 (before their first statement, `ig%epot = 0._dp`,
                 Pairs.F90:215,300,410,516):
                   if (gems_hip_on()) then
                     if (pair_hip_interact(ig, HIP_..., ig%e, ig%s, rs)) return
                   endif
  TB.F90         the same for smatb_interact (before TB.F90:465,
                 `call smatb_preinteraction(ig)`), with tb_hip_interact
  Integration.F90  a public function its_hip_kind(i): 1 when integrator i
                 steps with velocity_verlet_a/b, 2 with ermak_a/b, else 0
                 (the procedure pointers stepa/stepb, Integration.F90:37-38,
                 compared as integration_reversea compares them)
  Random.F90     a public subroutine lcg_hip_state(s, put, ok): the `prng lcg`
                 state (idum, iy, ix; Random.F90:33-35) out (put = .false.)
                 or in; ok = .false. when the stream is not the LCG
  Programs.f90   in dinamic, after its first interact(.false.)
                 (Programs.f90:129):
                   if (gems_hip_on()) then
                     if (hip_dinamic(steps, b_out, b_time)) return
                   endif
"""
import os
import re
import sys

PAIR_KINDS = [("lj", "HIP_LJ", "0._dp"), ("plj", "HIP_PLJ", "0._dp"),
              ("slj", "HIP_SLJ", "ig%rs"), ("wca", "HIP_WCA", "0._dp")]

ITS_KIND = """
integer function its_hip_kind(i)
! 1: integrator i steps with velocity_verlet_a/b; 2: ermak_a/b; 0: other.
use gems_constants, only: same_proc
use, intrinsic :: iso_c_binding
integer, intent(in) :: i
its_hip_kind = 0
if (.not. associated(its%o(i)%stepa) .or. .not. associated(its%o(i)%stepb)) return
if (same_proc(c_funloc(its%o(i)%stepa), c_funloc(velocity_verlet_a)) .and. &
    same_proc(c_funloc(its%o(i)%stepb), c_funloc(velocity_verlet_b))) its_hip_kind = 1
if (same_proc(c_funloc(its%o(i)%stepa), c_funloc(ermak_a)) .and. &
    same_proc(c_funloc(its%o(i)%stepb), c_funloc(ermak_b))) its_hip_kind = 2
end function its_hip_kind
"""

LCG_STATE = """
subroutine lcg_hip_state(s, put, ok)
! The `prng lcg` state (idum, iy, ix) out (put = .false.) or in.
use, intrinsic :: iso_c_binding
integer, intent(inout) :: s(3)
logical, intent(in)    :: put
logical, intent(out)   :: ok
ok = associated(ranu)
if (ok) ok = transfer(c_funloc(ranu), 0_c_intptr_t) == transfer(c_funloc(lcg_ranu), 0_c_intptr_t)
if (.not. ok) return
if (put) then
  idum = s(1)
  iy = s(2)
  ix = s(3)
else
  s = [idum, int(iy), int(ix)]
endif
end subroutine lcg_hip_state
"""


def once(text, pattern, repl, what, start=0):
    """Replace the first match of pattern at or after start; it must exist."""
    m = re.compile(pattern, re.M).search(text, start)
    if not m:
        sys.exit("patch_gems: %s not found" % what)
    return text[:m.start()] + repl(m) + text[m.end():], m.start()


def add_use(text, line):
    return once(text, r"^use gems_neighbour *$", lambda m: m.group(0) + "\n" + line,
                "`use gems_neighbour`")[0]


def hook(body):
    return "if (gems_hip_on()) then\n  %s\nendif\n" % body


def pairs(src):
    src = add_use(src, "use gems_plugin_hip, only: gems_hip_on, pair_hip_interact, "
                       "HIP_LJ, HIP_PLJ, HIP_SLJ, HIP_WCA")
    for k, tag, rs in PAIR_KINDS:
        _, at = once(src, r"^subroutine %s_interact\(ig\) *$" % k, lambda m: m.group(0),
                     "%s_interact" % k)
        src, _ = once(src, r"^ig%epot = 0\._dp *$",
                      lambda m: hook("if (pair_hip_interact(ig, %s, ig%%e, ig%%s, %s)) return"
                                     % (tag, rs)) + m.group(0),
                      "the first statement of %s_interact" % k, at)
    return src


def tb(src):
    src = add_use(src, "use gems_plugin_hip, only: gems_hip_on, tb_hip_interact")
    _, at = once(src, r"^subroutine smatb_interact\(ig\) *$", lambda m: m.group(0),
                 "smatb_interact")
    return once(src, r"^call smatb_preinteraction\(ig\) *$",
                lambda m: hook("if (tb_hip_interact(ig, ig%nz, ig%z, ig%ca, ig%eps, ig%p, "
                               "ig%q, ig%r0, ig%rci, ig%rce)) return") + m.group(0),
                "the first statement of smatb_interact", at)[0]


def integration(src):
    src = once(src, r"^public :: integration_stepb *$",
               lambda m: m.group(0) + "\npublic :: its_hip_kind", "public integration_stepb")[0]
    return once(src, r"^end module gems_integration", lambda m: ITS_KIND + "\n" + m.group(0),
                "end module gems_integration")[0]


def random(src):
    src = once(src, r"^public :: init_ran, free_ran, write_chpseed, read_chpseed *$",
               lambda m: m.group(0) + "\npublic :: lcg_hip_state", "public init_ran")[0]
    return once(src, r"^end module gems_random", lambda m: LCG_STATE + "\n" + m.group(0),
                "end module gems_random")[0]


def programs(src):
    _, at = once(src, r"^subroutine dinamic\(steps,b_out,b_time\) *$", lambda m: m.group(0),
                 "dinamic")
    src, _ = once(src, r"^use gems_input_parsing, only: load_blk, bloques, execute_block *$",
                  lambda m: m.group(0) + "\nuse gems_plugin_hip, only: gems_hip_on"
                                         "\nuse gems_plugin_hip_md, only: hip_dinamic",
                  "dinamic's use lines", at)
    return once(src, r"^call interact\(\.false\.\) *$",
                lambda m: m.group(0) + "\n" + hook("if (hip_dinamic(steps, b_out, b_time)) return"),
                "dinamic's first interact", at)[0]


PATCHES = {"Pairs.F90": pairs, "TB.F90": tb, "Integration.F90": integration,
           "Random.F90": random, "Programs.f90": programs}


def main():
    if len(sys.argv) != 3:
        sys.exit(__doc__)
    src_dir, out_dir = sys.argv[1], sys.argv[2]
    os.makedirs(out_dir, exist_ok=True)
    for name, fn in PATCHES.items():
        with open(os.path.join(src_dir, name), encoding="latin-1") as f:
            text = f.read()
        with open(os.path.join(out_dir, name), "w", encoding="latin-1") as f:
            f.write(fn(text))


if __name__ == "__main__":
    main()
