"""The Fortran interface module include/mbgpu.f90 against the C header, on the CPU.

The reference's host is Fortran.  This compiles the module with amdflang, links a small
Fortran program against libmbgpu.so (so every bind(c) name must resolve to an exported
symbol), and checks:
  * the bind(c) parameter types have the size and field offsets of mbgpu_elec_params and
    mbgpu_water_params as gcc lays them out from include/mbgpu.h;
  * the entries that need no GPU answer through the Fortran interfaces: the cutoff leaves,
    mbgpu_check_listttm2 on a good and a bad list, and mbgpu_create failing cleanly with a
    message when no GPU is present.
"""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOD = os.path.join(REPO, "include", "mbgpu.f90")
LIBDIR = os.path.join(REPO, "dlpoly2-mbpol-heat_amd")
FLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

pytestmark = pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang is not installed")

ELEC = ["rcut", "alpha", "kmax", "toler", "omega", "max_iter", "gamma", "q_o", "polr", "polr2", "athole", "athole12",
        "athole13", "ascc", "ascd", "compat_qd_column", "distribute_msite", "solver"]
WATER = ["rcut", "rvdw", "c6", "a6", "c8", "a8", "bond_key", "bond_prm", "ps_monomer", "dispersion", "long_range"]

F_PROG = """
program abi
  use iso_c_binding
  use mbgpu
  implicit none
  type(mbgpu_elec_params), target :: ep
  type(mbgpu_water_params), target :: wp
  integer(c_int) :: lst(6), bad(6)
  real(c_double) :: r
  type(c_ptr) :: ctx
  integer(c_intptr_t) :: b
  b = transfer(c_loc(ep), b)
  print '(a,i0)', 'elec size ', c_sizeof(ep)
{elec}
  b = transfer(c_loc(wp), b)
  print '(a,i0)', 'water size ', c_sizeof(wp)
{water}
  call mbpol_2b_cutoff(r)
  print '(a,f4.1)', 'cutoff2 ', r
  call mbpol_3b_cutoff(r)
  print '(a,f4.1)', 'cutoff3 ', r
  lst = [1, 2, 3, 4, 5, 6]
  bad = [1, 2, 3, 4, 6, 5]
  print '(a,i0)', 'check good ', mbgpu_check_listttm2(6, lst, 8)
  print '(a,i0)', 'check bad ', mbgpu_check_listttm2(6, bad, 8)
  ctx = mbgpu_create(2, 8)
  if (c_associated(ctx)) then
    print '(a)', 'create ok'
    call mbgpu_destroy(ctx)
  else
    print '(a,a)', 'create failed: ', mbgpu_error_text()
  end if
end program abi
"""

C_PROG = """
#include <stddef.h>
#include <stdio.h>
#include "mbgpu.h"
int main(void) {{
  printf("elec size %zu\\n", sizeof(mbgpu_elec_params));
{elec}
  printf("water size %zu\\n", sizeof(mbgpu_water_params));
{water}
  return 0;
}}
"""


def run(cmd, cwd, env=None):
    r = subprocess.run(cmd, cwd=cwd, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return r.stdout


@pytest.fixture(scope="module")
def outputs(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("fortran"))
    lib = os.path.join(LIBDIR, "libmbgpu.so")
    assert os.path.exists(lib), "build first: __graft_entry__.build()"
    fe = "\n".join(f"  print '(a,i0)', 'elec {f} ', transfer(c_loc(ep%{f}), b) - b" for f in ELEC)
    fw = "\n".join(f"  print '(a,i0)', 'water {f} ', transfer(c_loc(wp%{f}), b) - b" for f in WATER)
    with open(os.path.join(d, "abi.f90"), "w") as f:
        f.write(F_PROG.format(elec=fe, water=fw))
    ce = "\n".join(f'  printf("elec {f} %zu\\n", offsetof(mbgpu_elec_params, {f}));' for f in ELEC)
    cw = "\n".join(f'  printf("water {f} %zu\\n", offsetof(mbgpu_water_params, {f}));' for f in WATER)
    with open(os.path.join(d, "abi.c"), "w") as f:
        f.write(C_PROG.format(elec=ce, water=cw))
    run([FLANG, "-c", "-module-dir", d, MOD, "-o", "mbgpu_mod.o"], d)
    run([FLANG, "-I", d, "abi.f90", "mbgpu_mod.o", "-L" + LIBDIR, "-lmbgpu", "-Wl,-rpath," + LIBDIR, "-o", "abi_f"], d)
    run(["gcc", "-I", os.path.join(REPO, "include"), "abi.c", "-o", "abi_c"], d)
    env = dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
    fout = run([os.path.join(d, "abi_f")], d, env=env)
    cout = run([os.path.join(d, "abi_c")], d)
    return fout, cout


def parse(text):
    return {" ".join(line.split()[:-1]): line.split()[-1] for line in text.splitlines() if line.strip()}


def test_parameter_types_match_the_c_layout(outputs):
    fout, cout = outputs
    f, c = parse(fout), parse(cout)
    for k, v in c.items():
        assert f.get(k) == v, (k, f.get(k), v)


def test_entries_without_gpu_answer_through_the_module(outputs):
    f = outputs[0]
    lines = f.splitlines()
    assert "cutoff2  6.5" in lines and "cutoff3  4.5" in lines
    assert "check good 0" in lines and "check bad 2" in lines
    import torch
    if not torch.cuda.is_available():
        assert any(s.startswith("create failed: mbgpu_create: no HIP device") for s in lines), f
