"""Host-side logic of the library and the C ABI surface (CPU only: no
compute call reaches a device)."""
import glob
import os
import re

import numpy as np
import pytest

import mrlbm
from mrlbm import _lib, geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "bc_*.npz")))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(mrlbm_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding types every one of them
    assert sorted(set(syms) - set(_lib.exported_symbols())) == []


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_bc_classes_match_reference_fragments(path):
    """The kernels' boundary tables (known mask + inverse system built on the
    host) reproduce the reference's case formulas."""
    d = np.load(path)
    q = int(d["q"])
    for t, om, pop, out in zip(d["types"], d["omega"], d["pops"], d["out"]):
        got = mrlbm.bc_solve_host(q, int(t), float(om), out[1:4], pop)
        np.testing.assert_allclose(got, out, rtol=1e-11, atol=1e-13)


def test_bc_class_info_rejects_bad_lattice():
    with pytest.raises(mrlbm.MrlbmError):
        mrlbm.bc_class_info(15, 0x33, 1.0)


def brute_force_types(shape, walls, solid):
    nx, ny, nz = shape
    t = np.zeros(shape, np.uint32)
    n = (nx, ny, nz)
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                if solid[x, y, z]:
                    t[x, y, z] = 0xFF
                    continue
                bits = 0
                p = (x, y, z)
                for a in range(3):
                    if walls[a] and p[a] == 0:
                        bits |= geometry.octants_of(tuple(-1 if b == a else 0 for b in range(3)))
                    if walls[a] and p[a] == n[a] - 1:
                        bits |= geometry.octants_of(tuple(1 if b == a else 0 for b in range(3)))
                for dx in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dz in (-1, 0, 1):
                            if (dx, dy, dz) == (0, 0, 0):
                                continue
                            q = [x + dx, y + dy, z + dz]
                            ok = True
                            for a in range(3):
                                if q[a] < 0 or q[a] >= n[a]:
                                    if walls[a]:
                                        ok = False
                                    q[a] %= n[a]
                            if ok and solid[q[0], q[1], q[2]]:
                                bits |= geometry.octants_of((dx, dy, dz))
                t[x, y, z] = bits
    return t


def test_geometry_classify_matches_brute_force():
    rng = np.random.default_rng(3)
    shape = (7, 6, 5)
    solid = rng.random(shape) < 0.15
    for walls in ((0, 0, 0), (0, 1, 1), (1, 1, 1)):
        np.testing.assert_array_equal(geometry.classify(shape, walls, solid), brute_force_types(shape, walls, solid))


def test_octant_rule_of_reference_bc_id():
    # bc_id (src/lbmInitialization.cu:1080-1227): +x neighbour -> bits 1,3,5,7;
    # (+1,+1,0) -> 3,7; (+1,+1,+1) -> 7; (-1,-1,-1) -> 0
    assert geometry.octants_of((1, 0, 0)) == 0b10101010
    assert geometry.octants_of((1, 1, 0)) == 0b10001000
    assert geometry.octants_of((1, 1, 1)) == 0b10000000
    assert geometry.octants_of((-1, -1, -1)) == 0b00000001
    # the named faces of include/mrlbm_nodetype.h
    assert geometry.octants_of((0, 1, 0)) == 0xCC and geometry.octants_of((0, 0, -1)) == 0x0F


def test_cavity_lid_owns_edges():
    t = geometry.cavity(8, lid_axis=2)
    lid = t[:, :, 7]
    assert np.all(lid & 0x100)
    assert np.all((lid & 0xFF) & 0xF0 == 0xF0)
    # edge of the lid plane also carries the side wall's octants
    assert t[0, 3, 7] & 0xFF == (0xF0 | 0x55)
    assert t[3, 3, 3] == 0


def test_create_reports_errors_without_crashing(gpu_available):
    if gpu_available:
        pytest.skip("host without a GPU only")
    with pytest.raises(mrlbm.MrlbmError):
        mrlbm.Lattice((8, 8, 8))


def test_invalid_parameters_rejected():
    with pytest.raises(mrlbm.MrlbmError):
        mrlbm.Lattice((8, 8, 8), q=15)
    with pytest.raises(mrlbm.MrlbmError):
        mrlbm.Lattice((8, 8, 8), tau=0.5)
    with pytest.raises(mrlbm.MrlbmError):
        mrlbm.Lattice((12, 8, 8), tile=(8, 8, 8))


def test_duct_request_stream():
    from mrlbm import serve
    a = serve.duct_requests(400, rate=20.0, seed=11)
    b = serve.duct_requests(400, rate=20.0, seed=11)
    arr = np.array([r.arrival_s for r in a])
    assert np.all(np.diff(arr) > 0)
    assert abs(400 / arr[-1] - 20.0) < 3.0  # Poisson rate
    for r, s in zip(a, b):
        assert (r.nx, r.ny, r.nz, r.steps, r.nspheres, r.seed) == (s.nx, s.ny, s.nz, s.steps, s.nspheres, s.seed)
        assert r.nx % 16 == 0 and 64 <= r.nx <= 256 and r.ny == r.nz and r.ny % 16 == 0 and 32 <= r.ny <= 128
        assert 1 <= r.nspheres <= 8 and 500 <= r.steps <= 5000
        assert 1e-7 <= r.gx <= 1e-5 and 0.01 <= r.nu <= 0.1
        for k in range(r.nspheres):
            assert 3.0 <= r.sphere[k][3] <= max(3.0, r.ny / 6.0)


def test_read_voxel_csv(tmp_path):
    p = tmp_path / "v.csv"
    p.write_text("1,2,3\n0,0,0\n9,9,9\n4,1,2\n")
    s = geometry.read_voxel_csv(p, (5, 4, 4))
    assert s.sum() == 3 and s[1, 2, 3] and s[0, 0, 0] and s[4, 1, 2]


def test_spheres_voxelisation():
    s = geometry.spheres((16, 16, 16), [(8, 8, 8)], [3.0])
    x, y, z = np.meshgrid(*[np.arange(16)] * 3, indexing="ij")
    np.testing.assert_array_equal(s, (x - 8) ** 2 + (y - 8) ** 2 + (z - 8) ** 2 <= 9)


def test_fastdiv_exact(tmp_path):
    """The step kernels' tile coordinates use a multiply-high division by the
    tile counts (fastdiv.h): exact against the hardware division for every
    divisor 1..4096, large divisors and edge dividends (tests/fastdiv_check.cpp)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = tmp_path / "fastdiv_check"
    src = os.path.join(os.path.dirname(__file__), "fastdiv_check.cpp")
    subprocess.run([gxx, "-std=c++17", "-O2", src, "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout


def test_halo_chunk_tables(tmp_path):
    """The tile kernel's ghost-chunk tables move every halo cell once from
    the right ghost-plane cell and read only inside the owner's plane
    (host simulation of every lane, tests/halo_check.cpp)."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "halo_check"
    src = os.path.join(os.path.dirname(__file__), "halo_check.cpp")
    subprocess.run([hipcc, "-std=c++17", "-O1", "--offload-host-only", "-x", "hip", src, "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout
