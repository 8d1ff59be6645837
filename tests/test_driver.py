"""The host driver (sop_gpu.x) against the reference program on the same scripts.

CPU: `sop_gpu.x --check` parses a script and applies its load / set commands;
the derived state must equal what the reference's own read_input + set_params +
load produce (run in a subprocess: the reference keeps global state).
GPU: a full run of sop_gpu.x and of the reference program on the same script
and inputs writes the same update.out, trajectory, coordinate, velocity and
generator-state files.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from refs import Ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sop-cpu_amd", "sop_gpu.x")
GOLDEN = os.path.join(ROOT, "tests", "golden")
needs_ref = pytest.mark.skipif(not Ref.available(), reason="oracle/_ref not built")

BASE = """set dynamics underdamped;
set zeta 50;
set t_step 0.05;
set temp 0.6;
set cutofftype neighborlist;
set nnlup 50;
set nup 100;
set nstep {nstep};
load bonds bonds.dat;
load angles angles.dat;
load vdw vdw.dat;
load init init.dat;
run;
"""

SCRIPTS = {
    "standard": BASE.format(nstep=1000),
    "reordered": "set t_step 0.01; set zeta 3;\nset dynamics underdamped;   set temp 1.1;\n"
                 "set zeta 50; set run 5; set boxl 504; set ncell 42; set cutofftype celllist;\n"
                 "load bonds bonds.dat; load angles angles.dat;\nload vdw vdw.dat;\nload init init.dat;\n"
                 "set nstep 20;\nrun;\n",
    "overdamped": "set dynamics overdamped;\nset zeta 50;\nset temp 0.6;\nset nnlup 20;\nset nup 10;\n"
                  "set cutofftype neighborlist;\nset nstep 50;\nload bonds bonds.dat;\nload angles angles.dat;\n"
                  "load vdw vdw.dat;\nload init init.dat;\nrun;\n",
    "no_final_newline": BASE.format(nstep=200).rstrip("\n"),
}


def _ref_check(d, script):
    code = (
        "import sys, ctypes; sys.path.insert(0, %r); from refs import Ref; r = Ref();\n"
        "r.L.ref_read_input(b'input.txt'); r.L.ref_ex_cmds_norun()\n"
        "names = ['a1','a2','a3','a4','a5','h','halfh','zeta','T','boxl','ncell','lcell','nstep',"
        "'e_ang_ss_coeff','f_ang_ss_coeff']\n"
        "for n in names: print(n, repr(r.getd(n)))\n"
        "for n in ['nbead','nbnd','nang','ncon_att','ncon_rep','nup','nnlup','sim_type','neighborlist','celllist']:"
        " print(n, r.geti(n))\n"
        "print('ncmd', r.L.ref_ncmd())\n"
        "print('first_gasdev', repr(r.L.ref_gasdev()))\n" % os.path.join(ROOT, "tests"))
    out = subprocess.run([sys.executable, "-c", code], cwd=d, capture_output=True, text=True, check=True).stdout
    return {l.split()[0]: l.split()[1] for l in out.splitlines() if not l.startswith("[")}


def _drv_check(d):
    out = subprocess.run([DRIVER, "--check", "input.txt"], cwd=d, capture_output=True, text=True,
                         check=True).stdout
    return {l.split()[0]: l.split()[1] for l in out.splitlines() if not l.startswith("[")}


def _case_dir(tmp_path, name, inputs="n76"):
    d = tmp_path / name
    shutil.copytree(os.path.join(GOLDEN, inputs), d)
    with open(d / "input.txt", "w") as f:
        f.write(SCRIPTS[name])
    return str(d)


@needs_ref
@pytest.mark.parametrize("name", sorted(SCRIPTS))
def test_driver_parses_like_reference(tmp_path, name):
    d = _case_dir(tmp_path, name)
    ref = _ref_check(d, SCRIPTS[name])
    drv = _drv_check(d)
    for k, v in ref.items():
        if k in ("first_gasdev",) or "." in v or "e" in v or "inf" in v or "nan" in v:
            assert float(drv[k]) == float(v), (k, drv[k], v)
        else:
            assert int(drv[k]) == int(v), (k, drv[k], v)


def _run_reference(d):
    code = ("import sys; sys.path.insert(0, %r); from refs import Ref; r = Ref();"
            "r.L.ref_read_input(b'input.txt'); r.L.ref_ex_cmds()" % os.path.join(ROOT, "tests"))
    subprocess.run([sys.executable, "-c", code], cwd=d, capture_output=True, text=True, check=True)


def _read_lines(path):
    with open(path) as f:
        return [l.split() for l in f if l.strip()]


def _compare_text_numbers(a, b, rtol):
    la, lb = _read_lines(a), _read_lines(b)
    assert len(la) == len(lb), (a, len(la), len(lb))
    for ra, rb in zip(la, lb):
        assert len(ra) == len(rb)
        for x, y in zip(ra, rb):
            if x == y:
                continue
            fx, fy = float(x), float(y)
            # %f output: equal, or 1 unit in the 6th decimal from an energy
            # summed in another order (energies agree within 1e-10 relative)
            assert abs(fx - fy) <= max(rtol * abs(fy), 1.5e-6), (a, x, y)


@needs_ref
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["standard", "overdamped"])
def test_driver_run_matches_reference_outputs(tmp_path, name):
    dr = _case_dir(tmp_path / "ref", name)
    dd = _case_dir(tmp_path / "gpu", name)
    _run_reference(dr)
    print(f"[gpu] driver run {name}", flush=True)
    subprocess.run([DRIVER, "input.txt"], cwd=dd, capture_output=True, text=True, check=True, timeout=300)
    for f in ("traj.bin", "traj_uncorrected.bin", "coord.out", "unccoord.out", "veloc.out"):
        assert open(os.path.join(dr, f)).read() == open(os.path.join(dd, f)).read(), f
    _compare_text_numbers(os.path.join(dr, "update.out"), os.path.join(dd, "update.out"), 1e-10)
    assert open(os.path.join(dr, "ran_gen.dat")).read() == open(os.path.join(dd, "ran_gen.dat")).read()


RESTART = ("set restart on;\nset rgen_restart on;\nset istep_restart 200;\n"
           + BASE.format(nstep=400).replace("run;", "").rstrip("\n") + "\nrun;\n")
OUTPUTS = ("traj.bin", "traj_uncorrected.bin", "coord.out", "unccoord.out", "veloc.out", "ran_gen.dat")


def _run_driver(d, exe=DRIVER):
    return subprocess.run([exe, "input.txt"], cwd=d, capture_output=True, text=True, check=True, timeout=300)


def _same_outputs(dr, dd):
    for f in OUTPUTS:
        assert open(os.path.join(dr, f)).read() == open(os.path.join(dd, f)).read(), f
    _compare_text_numbers(os.path.join(dr, "update.out"), os.path.join(dd, "update.out"), 1e-10)


@needs_ref
@pytest.mark.gpu
def test_driver_file_restart_matches_reference(tmp_path):
    """A run to step 200, then a restart from its files (`set restart on; set rgen_restart on;
    set istep_restart 200`: coordinates, velocities and the generator state read back,
    sop.cpp:126-140, io.cpp:180-242, random_generator.cpp:51-70) to step 400: every output
    file equals the reference program's, including the appended update.out and trajectories."""
    dr = _case_dir(tmp_path / "ref", "standard")
    dd = _case_dir(tmp_path / "gpu", "standard")
    for d in (dr, dd):
        with open(os.path.join(d, "input.txt"), "w") as f:
            f.write(BASE.format(nstep=200))
    _run_reference(dr)
    print("[gpu] driver run to step 200", flush=True)
    _run_driver(dd)
    _same_outputs(dr, dd)
    for d in (dr, dd):
        with open(os.path.join(d, "input.txt"), "w") as f:
            f.write(RESTART)
    _run_reference(dr)
    print("[gpu] driver restart to step 400", flush=True)
    _run_driver(dd)
    _same_outputs(dr, dd)
    assert len(_read_lines(os.path.join(dd, "update.out"))) == 4


def _reference_final_state(d):
    """Run the reference program on d/input.txt and return its final pos, unc_pos, vel at full
    precision (read from its globals, not from its %f files)."""
    code = ("import sys, numpy as np; sys.path.insert(0, %r); from refs import Ref; r = Ref();"
            "r.L.ref_read_input(b'input.txt'); r.L.ref_ex_cmds();"
            "np.save('state.npy', np.stack([r.coords(0), r.coords(1), r.coords(2)]))" % os.path.join(ROOT, "tests"))
    subprocess.run([sys.executable, "-c", code], cwd=d, capture_output=True, text=True, check=True)
    return np.load(os.path.join(d, "state.npy"))


def _idx_xyz(path):
    return np.array([[float(t) for t in l.split()[1:4]] for l in open(path) if l.strip()]).ravel()


@needs_ref
@pytest.mark.gpu
def test_driver_lossless_restart_files(tmp_path):
    """`set restart_format lossless;` (a key the reference ignores) writes coordinates,
    velocities and the generator state at %.17g: the files hold the reference's final state
    bit for bit, and a restart from them -- by the reference and by the driver -- gives the
    same files again."""
    lossless = "set restart_format lossless;\n"
    dr = _case_dir(tmp_path / "ref", "standard")
    dd = _case_dir(tmp_path / "gpu", "standard")
    for d in (dr, dd):
        with open(os.path.join(d, "input.txt"), "w") as f:
            f.write(lossless + BASE.format(nstep=200))
    ref_state = _reference_final_state(dr)
    print("[gpu] driver run to step 200, lossless files", flush=True)
    _run_driver(dd)
    for k, f in enumerate(("coord.out", "unccoord.out", "veloc.out")):
        assert np.array_equal(_idx_xyz(os.path.join(dd, f)), ref_state[k]), f
    a = open(os.path.join(dr, "ran_gen.dat")).read().split()
    b = open(os.path.join(dd, "ran_gen.dat")).read().split()
    assert a[:2] + a[3:] == b[:2] + b[3:] and abs(float(a[2]) - float(b[2])) <= 5e-7
    # both programs restart from the driver's lossless files
    for f in OUTPUTS + ("update.out",):
        shutil.copy(os.path.join(dd, f), os.path.join(dr, f))
    for d in (dr, dd):
        with open(os.path.join(d, "input.txt"), "w") as f:
            f.write(RESTART)  # %f output again, as the reference writes
    _run_reference(dr)
    print("[gpu] driver restart from lossless files", flush=True)
    _run_driver(dd)
    _same_outputs(dr, dd)


ROUTE2 = os.path.join(ROOT, "sop-cpu_amd", "sop_ref_gpu.x")


@needs_ref
@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(ROUTE2), reason="sop_ref_gpu.x not built (needs the reference's sources)")
@pytest.mark.parametrize("name", ["standard", "overdamped"])
def test_reference_program_with_gpu_shim_matches_reference(tmp_path, name):
    """SOP-CPU's own program with the shim linked in place of its hot path (sop_ref_gpu.x,
    built by `make -C sop-cpu_amd route2`) writes the reference program's files."""
    dr = _case_dir(tmp_path / "ref", name)
    dd = _case_dir(tmp_path / "gpu", name)
    _run_reference(dr)
    print(f"[gpu] reference program with the GPU shim: {name}", flush=True)
    _run_driver(dd, ROUTE2)
    _same_outputs(dr, dd)
