"""Checkpoint file validation on the host (CPU only): files written here in
the layout include/mrlbm.h documents, then checked by
mrlbm_checkpoint_info, which reads no device. The GPU tests cover save and
load themselves."""
import struct

import numpy as np
import pytest

import mrlbm
from mrlbm import _lib

HDR = struct.Struct("<8sI11iIIq5d7Q")


def write_checkpoint(path, nx=16, ny=16, nz=8, q=19, tile=(8, 8, 8), storage=0, walls=(0, 1, 0), seed=3, step=40,
                     tau=0.6, rho0=1.0, force=(1e-5, 0, 0), mom_bytes=None, face_bytes=None, truncate=0,
                     magic=b"MRLBMCK2", version=2):
    bx, by, bz = tile
    ntiles = (nx // bx) * (ny // by) * (nz // bz)
    qf = 5 if q == 19 else 9
    sb = 4 if storage == 0 else 2
    mb = nx * ny * nz * 10 * sb if mom_bytes is None else mom_bytes
    fb = [by * bz * qf * ntiles * 4] * 2 + [bx * bz * qf * ntiles * 4] * 2 + [bx * by * qf * ntiles * 4] * 2
    if face_bytes is not None:
        fb = face_bytes
    hdr = HDR.pack(magic, version, nx, ny, nz, q, bx, by, bz, storage, *walls, seed, 0, step, tau, rho0, *force,
                   mb, *fb)
    rng = np.random.default_rng(0)
    payload = rng.bytes(nx * ny * nz * 10 * sb + sum(fb[:6]))
    data = hdr + payload
    if truncate:
        data = data[:-truncate]
    with open(path, "wb") as f:
        f.write(data)


def test_header_size_matches_the_documented_layout():
    assert HDR.size == 168  # MRLBM_CHECKPOINT_HEADER_BYTES


def test_valid_file_reports_its_parameters(tmp_path):
    p = tmp_path / "ck.bin"
    write_checkpoint(p, storage=mrlbm.STORE_FP16, seed=77, step=123, tau=0.57)
    params, step = mrlbm.checkpoint_info(p)
    assert step == 123
    assert (params["nx"], params["ny"], params["nz"], params["q"]) == (16, 16, 8, 19)
    assert (params["bx"], params["by"], params["bz"]) == (8, 8, 8)
    assert params["storage"] == mrlbm.STORE_FP16 and params["seed"] == 77
    assert params["tau"] == 0.57 and params["fx"] == 1e-5
    assert (params["wall_x"], params["wall_y"], params["wall_z"]) == (0, 1, 0)


@pytest.mark.parametrize("kind", ["truncated", "extra_bytes", "wrong_mom_bytes", "wrong_face_bytes", "magic",
                                  "version", "tile_not_dividing", "bad_q"])
def test_malformed_files_are_rejected(tmp_path, kind):
    p = tmp_path / "bad.bin"
    if kind == "truncated":
        write_checkpoint(p, truncate=7)
    elif kind == "extra_bytes":
        write_checkpoint(p)
        with open(p, "ab") as f:
            f.write(b"\0" * 4)
    elif kind == "wrong_mom_bytes":
        write_checkpoint(p, mom_bytes=16 * 16 * 8 * 10 * 4 - 4)
    elif kind == "wrong_face_bytes":
        write_checkpoint(p, face_bytes=[4096, 4096, 4096, 4096, 4096, 4092])
    elif kind == "magic":
        write_checkpoint(p, magic=b"MRLBMCK1")
    elif kind == "version":
        write_checkpoint(p, version=3)
    elif kind == "tile_not_dividing":
        write_checkpoint(p, nx=20)
    elif kind == "bad_q":
        write_checkpoint(p, q=15)
    with pytest.raises(mrlbm.MrlbmError) as e:
        mrlbm.checkpoint_info(p)
    assert e.value.code == -4  # MRLBM_EIO
    assert "checkpoint" in str(e.value)


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(mrlbm.MrlbmError) as e:
        mrlbm.checkpoint_info(tmp_path / "absent.bin")
    assert e.value.code == -4


def test_null_arguments_rejected():
    L = _lib.load()
    assert L.mrlbm_checkpoint_info(None, None, None) == -1
