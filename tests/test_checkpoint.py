"""Checkpoint files in the reference's HDF5 layout (IO/Writers.jl:41-134)
through libsfdio.so: "{prefix}-{istep}.h5" with datasets "re" and "im" of
the field's real type, dataspace (nz, ny, nx), x fastest.

The layout is checked with the HDF5 library's own tools where the image has
them (h5dump prints the dataspace and type and dumps the raw data; h5import
writes a file independently of this writer; h5repack rewrites ours chunked
and compressed), so that neither direction is checked only against itself."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

sio = pytest.importorskip("sfdhip.io")

TOOLS = "/opt/conda/bin"


def _tool(name):
    p = shutil.which(name) or os.path.join(TOOLS, name)
    if not os.path.exists(p):
        pytest.skip(f"{name} not available")
    return p


@pytest.fixture(autouse=True)
def _lib():
    if not os.path.exists(sio.IO_PATH):
        pytest.skip("libsfdio.so not built (no HDF5 headers at build time)")


def _state(shape, dtype, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("shape", [(16,), (24, 40), (6, 10, 8)])
def test_round_trip_is_exact(tmp_path, shape, dtype):
    a = _state(shape, dtype)
    p = str(tmp_path / "res-3.h5")
    sio.write_array(p, a)
    b = sio.read_array(p, shape, dtype)
    assert b.dtype == dtype and np.array_equal(a, b)


def test_read_converts_the_stored_type(tmp_path):
    """read! reads into Float64 whatever FT was stored (Writers.jl:127-131)."""
    a = _state((8, 12), np.complex64)
    p = str(tmp_path / "f32.h5")
    sio.write_array(p, a)
    assert np.array_equal(sio.read_array(p, (8, 12), np.complex128), a.astype(np.complex128))


def test_layout_seen_by_h5dump(tmp_path):
    h5dump = _tool("h5dump")
    a = _state((3, 4, 5), np.complex128, seed=2)   # nz=3, ny=4, nx=5
    p = str(tmp_path / "res-0.h5")
    sio.write_array(p, a)
    head = subprocess.run([h5dump, "-H", p], capture_output=True, text=True, check=True).stdout
    for name, ftype in (("re", "H5T_IEEE_F64LE"), ("im", "H5T_IEEE_F64LE")):
        block = head[head.index(f'DATASET "{name}"'):]
        assert ftype in block.split("}")[0]
        dims = re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(\s*([0-9, ]+)\)", block).group(1)
        assert [int(v) for v in dims.split(",")] == [3, 4, 5]
    for name, part in (("re", a.real), ("im", a.imag)):
        out = str(tmp_path / f"{name}.bin")
        subprocess.run([h5dump, "-d", f"/{name}", "-b", "LE", "-o", out, p], capture_output=True, check=True)
        raw = np.fromfile(out, dtype="<f8")
        assert np.array_equal(raw, part.ravel())  # C order of (nz, ny, nx): x fastest
    a32 = a.astype(np.complex64)
    sio.write_array(p, a32)
    head = subprocess.run([h5dump, "-H", p], capture_output=True, text=True, check=True).stdout
    assert head.count("H5T_IEEE_F32LE") == 2


def test_reads_a_file_written_by_h5import(tmp_path):
    h5import = _tool("h5import")
    a = _state((6, 10), np.complex128, seed=4)   # ny=6, nx=10
    args = []
    for name, part, size in (("re", a.real, 64), ("im", a.imag, 32)):
        binf = tmp_path / f"{name}.bin"
        np.ascontiguousarray(part).astype("<f8").tofile(binf)
        cfg = tmp_path / f"{name}.cfg"
        cfg.write_text(f"PATH {name}\nINPUT-CLASS FP\nINPUT-SIZE 64\nINPUT-BYTE-ORDER LE\nRANK 2\n"
                       f"DIMENSION-SIZES 6 10\nOUTPUT-CLASS FP\nOUTPUT-SIZE {size}\n"
                       f"OUTPUT-ARCHITECTURE IEEE\nOUTPUT-BYTE-ORDER LE\nCHUNKED-DIMENSION-SIZES 3 5\n"
                       f"COMPRESSION-TYPE GZIP\nCOMPRESSION-PARAM 6\n")
        args += [str(binf), "-c", str(cfg)]
    p = str(tmp_path / "res-7.h5")
    subprocess.run([h5import, *args, "-o", p], capture_output=True, check=True)
    b = sio.read_array(p, (6, 10), np.complex128)
    assert np.array_equal(b.real, a.real)
    assert np.array_equal(b.imag, a.imag.astype(np.float32).astype(np.float64))  # stored as Float32


def test_reads_our_file_after_h5repack(tmp_path):
    h5repack = _tool("h5repack")
    a = _state((4, 8, 16), np.complex64, seed=5)
    p, q = str(tmp_path / "a.h5"), str(tmp_path / "b.h5")
    sio.write_array(p, a)
    subprocess.run([h5repack, "-f", "GZIP=5", "-l", "CHUNK=2x4x8", p, q], capture_output=True, check=True)
    assert np.array_equal(sio.read_array(q, (4, 8, 16), np.complex64), a)


def test_errors(tmp_path):
    a = _state((8, 8), np.complex128)
    p = str(tmp_path / "x.h5")
    sio.write_array(p, a)
    with pytest.raises(sio.CheckpointError, match="dimension"):
        sio.read_array(p, (8, 16), np.complex128)
    with pytest.raises(sio.CheckpointError, match="rank"):
        sio.read_array(p, (2, 8, 8), np.complex128)
    with pytest.raises(sio.CheckpointError, match="cannot open"):
        sio.read_array(str(tmp_path / "missing.h5"), (8, 8))
    with pytest.raises(sio.CheckpointError, match="cannot create"):
        sio.write_array(str(tmp_path / "no-such-dir" / "y.h5"), a)
    assert sio.filename("res", 12) == "res-12.h5"
