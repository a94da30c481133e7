"""Checkpoint/resume for every storage format and for batch members, and
the field files (VTK, raw .bin, info) read back and checked against the
moments (through the C ABI)."""
import os
import struct

import numpy as np
import pytest

import mrlbm
from mrlbm import geometry

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def cavity(n=32, storage=mrlbm.STORE_FP32, seed=9, tau=0.6, q=19):
    lat = mrlbm.Lattice((n, n, n), q=q, tau=tau, walls=(1, 1, 1), storage=storage, seed=seed, tile=(8, 8, 8))
    lat.set_bc_velocity(0, (0.05, 0, 0))
    lat.set_node_types(geometry.cavity(n))
    lat.init_uniform()
    return lat


@pytest.mark.parametrize("storage", [mrlbm.STORE_FP16, mrlbm.STORE_BF16])
def test_checkpoint_resume_reduced_precision_bitwise(tmp_path, storage):
    """Save at step 40, resume in a fresh context and run 60 more: the
    FP16/BF16 state (stored words and the stochastic-rounding key: node,
    step, seed) continues exactly as the straight 100-step run."""
    say("checkpoint storage", storage)
    with cavity(storage=storage) as a:
        a.step(100)
        straight = a.moments()
    with cavity(storage=storage) as b:
        b.step(40)
        b.checkpoint_save(tmp_path / "ck.bin")
    params, step = mrlbm.checkpoint_info(tmp_path / "ck.bin")
    assert step == 40 and params["storage"] == storage and params["seed"] == 9
    with cavity(storage=storage) as c:
        c.checkpoint_load(tmp_path / "ck.bin")
        assert c.info()["step"] == 40
        c.step(60)
        assert np.array_equal(c.moments(), straight)


def test_checkpoint_resume_batch_member_bitwise(tmp_path):
    """A lattice saved from a batch and resumed into another batch continues
    as the straight batched run (and as the lattice alone)."""
    steps_a, steps_b = 30, 50
    with cavity(32, storage=mrlbm.STORE_FP16, seed=1) as lone:
        lone.step(steps_a + steps_b)
        straight = lone.moments()
    with mrlbm.Batch(q=19, tile=(8, 8, 8), storage=mrlbm.STORE_FP16) as b:
        lats = [cavity(32, storage=mrlbm.STORE_FP16, seed=1), cavity(16, storage=mrlbm.STORE_FP16, seed=2)]
        for lat in lats:
            b.add(lat)
        b.step(steps_a, graph_steps=10)
        lats[0].checkpoint_save(tmp_path / "m.bin")
        for lat in lats:
            lat.close()
    with mrlbm.Batch(q=19, tile=(8, 8, 8), storage=mrlbm.STORE_FP16) as b:
        other = cavity(16, storage=mrlbm.STORE_FP16, seed=2)
        res = cavity(32, storage=mrlbm.STORE_FP16, seed=1)
        res.checkpoint_load(tmp_path / "m.bin")
        b.add(other)
        b.add(res)
        b.step(steps_b, graph_steps=10)
        assert res.info()["step"] == steps_a + steps_b
        assert np.array_equal(res.moments(), straight)
        other.close()
        res.close()


@pytest.mark.parametrize("field,value", [("tau", 0.61), ("seed", 10), ("storage", mrlbm.STORE_BF16)])
def test_checkpoint_load_rejects_mismatch(tmp_path, field, value):
    with cavity(16, storage=mrlbm.STORE_FP16) as a:
        a.step(5)
        a.checkpoint_save(tmp_path / "ck.bin")
        before = a.moments()
    kw = {"storage": mrlbm.STORE_FP16, "seed": 9, "tau": 0.6}
    kw[field] = value
    with cavity(16, **kw) as b:
        b.step(3)
        m0 = b.moments()
        with pytest.raises(mrlbm.MrlbmError) as e:
            b.checkpoint_load(tmp_path / "ck.bin")
        say(e.value)
        assert b.info()["step"] == 3  # untouched
        assert np.array_equal(b.moments(), m0)
    assert before.shape == m0.shape


def test_checkpoint_load_truncated_file_leaves_state(tmp_path):
    with cavity(16) as a:
        a.step(5)
        a.checkpoint_save(tmp_path / "ck.bin")
    data = (tmp_path / "ck.bin").read_bytes()
    (tmp_path / "cut.bin").write_bytes(data[:-100])
    with cavity(16) as b:
        b.step(2)
        m0 = b.moments()
        with pytest.raises(mrlbm.MrlbmError):
            b.checkpoint_load(tmp_path / "cut.bin")
        assert np.array_equal(b.moments(), m0) and b.info()["step"] == 2


def read_vtk(path):
    """Parse the legacy-VTK STRUCTURED_POINTS file: dims, rho and velocity
    (big-endian float32 payloads)."""
    data = open(path, "rb").read()
    i = data.index(b"LOOKUP_TABLE default\n") + len(b"LOOKUP_TABLE default\n")
    head = data[:i].decode()
    dims = [int(v) for v in head.split("DIMENSIONS")[1].split("\n")[0].split()]
    n = dims[0] * dims[1] * dims[2]
    rho = np.frombuffer(data, ">f4", n, i)
    j = data.index(b"VECTORS velocity float\n", i + 4 * n) + len(b"VECTORS velocity float\n")
    vel = np.frombuffer(data, ">f4", 3 * n, j).reshape(n, 3)
    assert j + 12 * n == len(data)
    return dims, rho, vel, head


def test_vtk_values_and_byte_order(tmp_path):
    shape = (16, 8, 24)
    with mrlbm.Lattice(shape, q=19, tau=0.6) as lat:
        lat.init_taylor_green(0.05)
        lat.step(3)
        m = lat.moments()
        lat.save_vtk(tmp_path / "f.vtk")
    dims, rho, vel, head = read_vtk(tmp_path / "f.vtk")
    assert head.startswith("# vtk DataFile Version 3.0\n") and "BINARY" in head
    assert dims == list(shape)
    np.testing.assert_array_equal(rho, (1.0 + m[0]).astype(np.float32))
    np.testing.assert_array_equal(vel, (m[1:4] / 3.0).T.astype(np.float32))


def test_bin_and_info_files_for_post_tools(tmp_path):
    shape = (16, 8, 24)
    with mrlbm.Lattice(shape, q=19, tau=0.6, force=(1e-6, 0, 0)) as lat:
        lat.init_taylor_green(0.05)
        lat.step(250)
        m = lat.moments()
        lat.save_bin(tmp_path, "007")
        lat.save_info(tmp_path, "007", case_name="taylorGreen", u_max=0.05, save_steps=250, mlups=1234.5)
    names = sorted(os.listdir(tmp_path))
    assert names == ["007_info.txt", "007_rho000250.bin", "007_ux000250.bin", "007_uy000250.bin",
                     "007_uz000250.bin"]
    # post/fileTreat.py reads float32 arrays reshaped (NZ, NY, NX) in C order
    rho = np.fromfile(tmp_path / "007_rho000250.bin", np.float32)
    ux = np.fromfile(tmp_path / "007_ux000250.bin", np.float32)
    np.testing.assert_array_equal(rho, (1.0 + m[0]).astype(np.float32))
    np.testing.assert_array_equal(ux, m[1].astype(np.float32))  # the stored 3u, as saveMacr writes it
    assert rho.reshape(shape[2], shape[1], shape[0]).shape == (24, 8, 16)
    info = {}
    for line in open(tmp_path / "007_info.txt"):
        if ":" in line:
            k, v = line.split(":", 1)
            info.setdefault(k.strip(), v.strip())
    assert info["Simulation ID"] == "007" and info["Precision"] == "float"
    assert (int(info["NX"]), int(info["NY"]), int(info["NZ"]), int(info["NZ_TOTAL"])) == (16, 8, 24, 24)
    assert float(info["Tau"]) == 0.6 and float(info["Umax"]) == 0.05 and float(info["FX"]) == 1e-6
    assert int(info["Nsteps"]) == 250 and info["BC type"] == "taylorGreen"
