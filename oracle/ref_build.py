#!/usr/bin/env python3
"""Build recipe for the reference's own per-timestep path (test
infrastructure only; run in the build container, where the reference
sources lie under /root/reference).

For each case it
  * writes an edited copy of the reference's configuration template
    Scripts/parameters.F90_template, with the case's values substituted
    into the template's own parameter lines (the reference configures a
    run exactly this way: install.sh copies the template to
    program/parameters.F90 and the user edits it);
  * reaches Scripts/parallel.h_template (serial: _Nth = _Nr = 1) through
    a symlink named parallel.h;
  * compiles, unmodified and where they lie, the reference modules of the
    hot path (modules/dyn_mpi, meshs, legendre, variables, transform.fftw3,
    timestep, velocity, magnetic, codensity, composition, rotation,
    nonlinear) and the auxiliary Legendre routines (auxiliary/*.f) with
    amdflang, plus the driver oracle/ref_step.F90;
  * links the FFTW3 and LAPACK entry points of the image's MKL
    (/opt/conda/lib/libmkl_rt.so), the libraries the reference is built
    against (Scripts/Makefile_template:64-77).
Only the executable oracle/_ref/<case>/ref_step is kept, with a gzip copy
ref_step.gz (the executables hold the reference's zero-initialised module
arrays as data, 1.3 GB at L = M = 170, N = 96; the copy is what travels to
other machines, see ref_io.exe); the edited template, objects and module
files are deleted after the link.  Nothing of the reference is copied into
the repository.

With --dropin it builds instead oracle/_ref/<case>/dropin_step: the same
driver and the same reference modules, except that modules/legendre.F90 and
modules/transform.fftw3.F90 are replaced by the repository's drop-in
modules (leeds-code_amd/fortran/legendre.hip.F90, transform.hip.F90, with
lsd_c.F90 and lsd_dropin.F90), linked against leeds-code_amd/liblsdgpu.so:
the reference's own physics modules stepping with the Legendre tables and
transforms on the GPU.  Run it on a GPU machine.

Usage: python oracle/ref_build.py [--dropin] [case ...]      (default: every case)
"""
import os
import gzip
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = os.environ.get("REF", "/root/reference")
FC = os.environ.get("REF_FC", "/opt/rocm/bin/amdflang")
MKL = os.environ.get("MKL_DIR", "/opt/conda/lib")
sys.path.insert(0, os.path.join(ROOT, "leeds-code_amd"))

MODULES = ["dyn_mpi", "meshs", "legendre", "variables", "transform.fftw3", "timestep", "velocity",
           "magnetic", "codensity", "composition", "rotation", "nonlinear"]
AUX = ["gauwts", "schnla", "pmm", "pmm1", "plm", "dpmm", "dpmm1", "dplm"]


def cases():
    """Case name -> lsd_config of the run (the same values the GPU path and
    the C oracle are given)."""
    from lsd.config import config1, config3, config_small, christensen_case1, christensen_case2
    return {"cfg1": config1(), "small": config_small(), "cfg3": config3(),
            "adapt": config_small(dt=-1e-4), "adapt1": config_small(dt=-1e-4, step_control=1),
            "case1": christensen_case1(),
            "ic": config_small(Nic=8, rot=1), "case2": christensen_case2(),
            "sf2": config_small(vel_bc=2), "sf3": config_small(vel_bc=3), "sf4": config_small(vel_bc=4)}


def fdouble(x):
    """A Fortran double-precision literal that reads back to x exactly."""
    s = "%.17e" % float(x)
    return s.replace("e", "d")


def flogical(b):
    return ".true." if b else ".false."


def edits(cfg):
    """(parameter name, value text) pairs for the template's lines."""
    H, Th, Ph, Th2 = cfg.sizes()
    ed = [("i_N", str(cfg.N)), ("i_Nic", str(max(1, cfg.Nic))), ("i_L", str(cfg.L)), ("i_M", str(cfg.M)), ("i_Mp", str(cfg.Mp)),
          ("d_Ra", fdouble(cfg.Ra)), ("d_Ra_comp", fdouble(cfg.Ra_comp)), ("d_Pr", fdouble(cfg.Pr)),
          ("d_Sc", fdouble(cfg.Sc)), ("d_Pm", fdouble(cfg.Pm)), ("d_E", fdouble(cfg.E)),
          ("d_rratio", fdouble(cfg.rratio)),
          ("b_vel_tstep", flogical(cfg.vel)), ("b_cod_tstep", flogical(cfg.cod)),
          ("b_comp_tstep", flogical(cfg.comp)), ("b_mag_tstep", flogical(cfg.mag)),
          ("i_vel_bc", str(cfg.vel_bc)), ("i_cod_bc", str(cfg.cod_bc)), ("i_comp_bc", str(cfg.comp_bc)),
          ("d_source", fdouble(cfg.source)), ("d_source_comp", fdouble(cfg.source_comp)),
          ("i_time_step_control", "1" if cfg.step_control == 1 else "2"), ("d_dterr", fdouble(cfg.dterr)),
          ("d_timestep", fdouble(cfg.dt)), ("d_implicit", fdouble(cfg.implicit)),
          ("b_rot_tstep", flogical(cfg.rot))]
    # grids that differ from the template's 3/2 rule: the rule lines
    # themselves take the case's sizes (Ph = 2 i_Ma, so i_Ma = Ph/2)
    if Th != cfg.L * 3 // 2:
        ed.append(("i_Th", str(Th)))
    if Ph != (cfg.M * 3 // 2) * 2:
        ed.append(("i_Ma", str(Ph // 2)))
    return ed


def edited_template(cfg):
    with open(os.path.join(REF, "Scripts", "parameters.F90_template")) as f:
        text = f.read()
    for name, value in edits(cfg):
        pat = re.compile(r"(parameter\s*::\s*%s\s*=\s*)([^!\n]*?)(\s*(!.*)?)$" % re.escape(name), re.M)
        text, n = pat.subn(lambda m: m.group(1) + value + m.group(3), text, count=1)
        if n != 1:
            raise RuntimeError("template line for %s not found" % name)
    return text


def run(cmd, cwd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


DROPIN = ["lsd_c", "lsd_dropin", "legendre.hip", "transform.hip"]
LIBDIR = os.path.join(ROOT, "leeds-code_amd")


def build(case, cfg, dropin=False):
    obj = os.path.join(HERE, "_ref", "_obj", case + ("_dropin" if dropin else ""))
    out = os.path.join(HERE, "_ref", case)
    shutil.rmtree(obj, ignore_errors=True)
    os.makedirs(os.path.join(obj, "inc", "sub"))
    os.makedirs(out, exist_ok=True)
    os.symlink(os.path.join(REF, "Scripts", "parallel.h_template"), os.path.join(obj, "inc", "parallel.h"))
    with open(os.path.join(obj, "parameters.F90"), "w") as f:
        f.write(edited_template(cfg))
    big = ["-mcmodel=medium"] if cfg.sizes()[1] * cfg.sizes()[2] * cfg.N > 1 << 22 else []
    flags = ["-cpp", "-O2", "-ffp-contract=off", "-I" + os.path.join(obj, "inc", "sub"),
             "-I" + os.path.join(obj, "inc")] + big
    srcs = [os.path.join(obj, "parameters.F90")]
    for m in MODULES:
        if dropin and m == "legendre":
            # the binding and the shared context first, then the drop-ins in
            # the places of the modules they replace
            srcs += [os.path.join(LIBDIR, "fortran", d + ".F90") for d in DROPIN[:3]]
        elif dropin and m == "transform.fftw3":
            srcs.append(os.path.join(LIBDIR, "fortran", "transform.hip.F90"))
        else:
            srcs.append(os.path.join(REF, "modules", m + ".F90"))
    objs = []
    for s in srcs:
        o = os.path.join(obj, os.path.basename(s).replace(".F90", ".o"))
        run([FC] + flags + ["-c", s, "-o", o], obj)
        objs.append(o)
    for a in AUX:
        o = os.path.join(obj, a + ".o")
        run([FC, "-O2", "-ffp-contract=off"] + big + ["-c", os.path.join(REF, "auxiliary", a + ".f"), "-o", o], obj)
        objs.append(o)
    drv = os.path.join(obj, "ref_step.o")
    run([FC] + flags + ["-c", os.path.join(HERE, "ref_step.F90"), "-o", drv], obj)
    exe = os.path.join(out, "dropin_step" if dropin else "ref_step")
    libs = ["-L" + MKL, "-lmkl_rt", "-Wl,-rpath," + MKL]
    if dropin:
        # liblsdgpu.so found relative to the executable, oracle/_ref/<case>/
        libs += ["-L" + LIBDIR, "-llsdgpu", "-Wl,-rpath,$ORIGIN/../../../leeds-code_amd"]
    run([FC] + big + ["-o", exe, drv] + objs + libs, obj)
    shutil.rmtree(obj)
    with open(exe, "rb") as f, gzip.open(exe + ".gz", "wb", compresslevel=6) as g:
        shutil.copyfileobj(f, g, 1 << 24)
    return exe


def main(argv):
    if not os.path.isdir(os.path.join(REF, "modules")):
        print("reference sources not present; nothing to build")
        return 0
    dropin = "--dropin" in argv
    argv = [a for a in argv if a != "--dropin"]
    allc = cases()
    names = argv or list(allc)
    for n in names:
        build(n, allc[n], dropin)
    shutil.rmtree(os.path.join(HERE, "_ref", "_obj"), ignore_errors=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
