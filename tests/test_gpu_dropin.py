"""Drop-in proof: the reference's own MPAS and MOM6 drivers (code.f90, mom6_code.f90), compiled
unchanged against the reference's modules and against the sal-fort_amd Fortran modules (oracle
Makefile target `drivers`), run on the same input files; the outputs agree within 1e-12.

The drivers read ./namelist.nml and ./data-files/2313486_mpas_grid_{areas,x,y,z,ssh}.csv and write
./run-output/2313486_sal_{x,y}.csv (code.f90:26-56, :107-114; mom6_code.f90:77-99, :340-356)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu

REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def _setup(tmp, n, thresh, degree, theta=0.7, seed=3):
    import salfort as S
    os.makedirs(os.path.join(tmp, "data-files"), exist_ok=True)
    os.makedirs(os.path.join(tmp, "run-output"), exist_ok=True)
    x, y, z, a = S.gen_mpas(n, seed=seed)
    s = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=seed + 1)
    for name, v in [("areas", a), ("x", x), ("y", y), ("z", z), ("ssh", s)]:
        with open(os.path.join(tmp, "data-files", f"2313486_mpas_grid_{name}.csv"), "w") as f:
            f.write("".join(repr(float(t)) + "\n" for t in v))
    with open(os.path.join(tmp, "namelist.nml"), "w") as f:
        f.write(f"&params\ntheta={theta}\npoint_count={n}\ninterp_degree={degree}\ncluster_thresh={thresh}\n/\n")


MPIEXEC = "/opt/conda/bin/mpiexec"


def _run(exe, tmp, ranks=1):
    cmd = [exe] if ranks == 1 else [MPIEXEC, "-n", str(ranks), exe]
    r = subprocess.run(cmd, cwd=tmp, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    print(f"ran {os.path.basename(exe)} on {ranks} rank(s)", flush=True)
    gx = np.loadtxt(os.path.join(tmp, "run-output", "2313486_sal_x.csv"))
    gy = np.loadtxt(os.path.join(tmp, "run-output", "2313486_sal_y.csv"))
    return gx, gy, r.stdout


@pytest.mark.parametrize("driver", ["mpas", "mom6"])
def test_reference_driver_unchanged_on_dropin(cuda, tmp_path, driver):
    ref_exe = os.path.join(REF_DIR, f"{driver}_ref")
    new_exe = os.path.join(REF_DIR, f"{driver}_salfort")
    if not (os.path.exists(ref_exe) and os.path.exists(new_exe)):
        pytest.skip("reference drivers not built (make -C oracle drivers, in the build container)")
    a, b = str(tmp_path / "ref"), str(tmp_path / "salfort")
    for d in (a, b):
        _setup(d, 20000, 33, 4)
    rx, ry, rout = _run(ref_exe, a)
    sx, sy, sout = _run(new_exe, b)
    assert rel_l2(sx, sy, rx, ry) < 1e-12
    # the drivers' own counts match: panels and interactions
    for key in ("total cube panels", "total interactions"):
        rl = [l for l in rout.splitlines() if key in l]
        sl = [l for l in sout.splitlines() if key in l]
        assert rl == sl, (rl, sl)


# the reference's own MOM6 driver ends in a segmentation fault at 3 ranks
@pytest.mark.parametrize("driver,ranks", [("mom6", 2), ("mom6", 4), ("mom6", 8), ("mpas", 2)])
def test_driver_at_several_ranks(cuda, tmp_path, driver, ranks):
    """The drivers at several MPI ranks (all on the one GPU).  mom6_code.f90: each rank builds the
    tree on the whole grid's coordinates, its own single-tree lists, its share of the proxy weights
    (added up by the driver's Sum_Array), and its PP sums over own points and halo copies
    (unowned_sources / relabel_copy_unowned / Communicate_sshs).  code.f90: fast_sum evaluates the
    whole list on rank 0 and adds nothing on the others, so the driver's Sum_Array gives the
    reference's totals.  Same output as the reference's driver at the same rank count (the MOM6
    output order is rank by rank), within 1e-12."""
    ref_exe = os.path.join(REF_DIR, f"{driver}_ref")
    new_exe = os.path.join(REF_DIR, f"{driver}_salfort")
    if not (os.path.exists(ref_exe) and os.path.exists(new_exe) and os.path.exists(MPIEXEC)):
        pytest.skip("reference drivers or mpiexec not available")
    a, b = str(tmp_path / "ref"), str(tmp_path / "salfort")
    for d in (a, b):
        _setup(d, 20000, 33, 4)
    rx, ry, rout = _run(ref_exe, a, ranks)
    sx, sy, sout = _run(new_exe, b, ranks)
    assert rel_l2(sx, sy, rx, ry) < 1e-12
    # (the per-rank "Rank: r interactions: k" lines differ by design: rank 0 evaluates the whole list)
    for key in ("total cube panels", "total interactions", "unowned points"):
        rl = sorted(l.strip() for l in rout.splitlines() if key in l)
        sl = sorted(l.strip() for l in sout.splitlines() if key in l)
        assert rl == sl, (rl, sl)


def _write_grid(path, x, y, z, area, ssh, thresh, degree, theta):
    with open(path, "wb") as f:
        np.array([len(x), thresh, degree], dtype=np.int32).tofile(f)
        np.array([theta], dtype=np.float64).tofile(f)
        for v in (x, y, z, area, ssh):
            np.ascontiguousarray(v, dtype=np.float64).tofile(f)


def _run_interact(exe, tmp, n):
    r = subprocess.run([exe, os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")], cwd=tmp,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    raw = open(os.path.join(tmp, "out.bin"), "rb").read()
    k = 8 if "mom6" in exe else 4
    counts = np.frombuffer(raw[:k], dtype=np.int32)
    v = np.frombuffer(raw[k:], dtype=np.float64).reshape(4, n)
    return counts, v


@pytest.mark.parametrize("thresh,degree", [(33, 4), (65, 7)])
def test_mpas_per_interaction_routines(cuda, oracle, tmp_path, thresh, degree):
    """pp_interact / pc_interact / cp_interact / cc_interact of the drop-in fast_sum_module, called
    for every list entry one at a time as the reference's fast_sum dispatches them
    (fast_sum_module.f90:303-329): the totals equal the drop-in fast_sum's and the oracle's within
    1e-12 (every entry type occurs)."""
    import salfort as S
    exe = os.path.join(REF_DIR, "mpas_interact")
    if not os.path.exists(exe):
        pytest.skip("per-interaction test programs not built (make -C oracle interact)")
    n, theta = 6000, 0.7
    x, y, z, area = S.gen_mpas(n, seed=12)
    ssh = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=13)
    _write_grid(str(tmp_path / "in.bin"), x, y, z, area, ssh, thresh, degree, theta)
    (entries,), (gx, gy, fx, fy) = _run_interact(exe, str(tmp_path), n)
    oracle.tree_build(x, y, z, thresh)
    _, _, ty = oracle.dtt(theta, thresh)
    assert entries == len(ty) and all((ty == k).any() for k in range(4))
    ox, oy = oracle.fast_sum(x, y, z, area, ssh, degree)
    assert rel_l2(gx, gy, fx, fy) < 1e-12
    assert rel_l2(gx, gy, ox, oy) < 1e-12


def test_mom6_per_interaction_routines(cuda, oracle, tmp_path):
    """pc_interact (each entry with its panel's slice of the proxy weights) and pp_interact of the
    drop-in mom6_fast_sum_module, called for every single-tree entry one at a time as pc_compute and
    pp_compute loop them (mom6_fast_sum_module.f90:153-192): the totals equal the drop-in list
    routines' and the oracle's MOM6 sums (with the reference's transposed weight read) within 1e-12."""
    import salfort as S
    exe = os.path.join(REF_DIR, "mom6_interact")
    if not os.path.exists(exe):
        pytest.skip("per-interaction test programs not built (make -C oracle interact)")
    n, thresh, degree, theta = 6000, 33, 4, 0.7
    x, y, z, area = S.gen_mpas(n, seed=14)
    ssh = S.gen_ssh(x, y, z, lmax=16, amp=0.1, noise=0.02, seed=15)
    _write_grid(str(tmp_path / "in.bin"), x, y, z, area, ssh, thresh, degree, theta)
    (npc, npp), (gx, gy, fx, fy) = _run_interact(exe, str(tmp_path), n)
    assert npc > 1000 and npp > 1000
    oracle.tree_build(x, y, z, thresh)
    oracle.stt(theta, thresh, x, y, z)
    g = S.Geometry(x, y, z, thresh)
    pp = oracle.assign(x, y, z, g.counts()[1] + 1)
    w = oracle.proxy_source(pp, x, y, z, ssh, area, degree)
    ox, oy = oracle.mom6_sums(x, y, z, area, ssh, degree, w)
    assert rel_l2(gx, gy, fx, fy) < 1e-12
    assert rel_l2(gx, gy, ox, oy) < 1e-12
