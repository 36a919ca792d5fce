"""The v2 input files and step buffers (awp_amd/inputs.py) against the
known answers of the reference's own tests (tests/readers/test_input.c,
tests/buffers/test_buffer.cu) and its fixture files (tests/fixtures/, copied
as data under tests/golden/fixtures/)."""
import shutil
from pathlib import Path

import numpy as np
import pytest

from awp_amd import inputs as I

FIX = Path(__file__).resolve().parent / "golden" / "fixtures"


@pytest.fixture
def workdir(tmp_path, monkeypatch):
    """Input files name data files relative to the working directory."""
    for f in FIX.iterdir():
        shutil.copy(f, tmp_path / f.name)
    monkeypatch.chdir(tmp_path)
    return tmp_path


def test_parse_arg_known_answers():
    # test_input.c:171-190
    assert I.parse_arg("a:b")[0] == I.ERR_CONFIG_PARSE_ARG
    assert I.parse_arg("=b")[0] == I.ERR_CONFIG_PARSE_ARG
    assert I.parse_arg("a=")[0] == I.ERR_CONFIG_PARSE_ARG
    assert I.parse_arg("test=a") == (I.SUCCESS, "test", "a")


def test_parse_known_answers():
    # test_input.c:191-215
    out = I.Input()
    for line in ("file=test", "length=2", "stride=10", "degree=3", "steps=1e3", "num_writes=10",
                 "gpu_buffer_size=10", "cpu_buffer_size=100"):
        assert out.parse(line) == I.SUCCESS
    assert (out.file, out.length, out.stride, out.degree, out.steps, out.num_writes,
            out.gpu_buffer_size, out.cpu_buffer_size) == ("test", 2, 10, 3, 1000, 10, 10, 100)
    assert out.parse("colour=3") == I.ERR_CONFIG_PARSE_UNKNOWN_ARG


def test_init_and_header_of_input1(workdir):
    # test_input.c:62-73, 120-131
    err, inp = I.input_init("input1.txt")
    assert err == I.SUCCESS
    assert (str(inp.version), inp.file, inp.length, inp.steps, inp.stride, inp.degree,
            inp.gpu_buffer_size, inp.cpu_buffer_size) == ("2.0.0", "output.bin", 4, 1000, 10, 3,
                                                          10, 10)
    assert np.array_equal(inp.x, [0.0, 1.5, 0.0, 1.5])
    assert np.array_equal(inp.y, [0.0, 0.0, 1.5, 1.5])
    assert np.array_equal(inp.type, [0, 0, 0, 0])
    assert inp.check_header() == I.SUCCESS
    inp.steps = 101
    assert inp.check_header() == I.ERR_CONFIG_PARSE_NOT_DIVISIBLE
    assert not (workdir / "output.bin").exists()   # the writeability probe cleans up


def test_file_readable_and_writeable(workdir):
    # test_input.c:75-118
    (workdir / "tmp").write_text("")
    assert I.file_readable("tmp") == I.SUCCESS
    (workdir / "tmp").unlink()
    assert I.file_readable("tmp") == I.ERR_FILE_READ
    assert I.file_writeable("tmp") == I.SUCCESS
    assert not (workdir / "tmp").exists()
    assert I.file_writeable("tmp/tmp") == I.ERR_FILE_WRITE


def test_write_read_round_trip(workdir):
    # test_input.c:133-152
    err, out = I.input_init("input1.txt")
    assert err == I.SUCCESS
    assert out.write("input.txt") == I.SUCCESS
    err, back = I.input_init("input.txt")
    assert err == I.SUCCESS
    assert back.equals(out)
    assert np.array_equal(back.x, out.x) and np.array_equal(back.z, out.z)


def test_fixture_files(workdir):
    err, rec = I.input_init("receiver.txt")
    assert err == I.SUCCESS
    assert (rec.length, rec.steps, rec.degree, rec.gpu_buffer_size, rec.cpu_buffer_size,
            rec.num_writes) == (4, 600, 3, 2, 3, 1)
    assert np.array_equal(rec.x, [0.0, 0.5, 0.0, 0.5])
    # length=4 with five rows: the fifth is not read
    err, src = I.input_init("source.txt")
    assert err == I.SUCCESS and src.length == 4 and src.degree == 1
    assert np.array_equal(src.x, [0.5, 0.5, 1.5, 0.0])
    # 2.1.0 is compatible; comment lines are skipped
    err, dm = I.input_init("source_dm.txt")
    assert err == I.SUCCESS and str(dm.version) == "2.1.0" and dm.length == 3
    assert np.array_equal(dm.z, [-2.0, -11.0, -34.0])
    # a version-1 file is refused
    assert I.input_init("recvs.in")[0] == I.ERR_WRONG_VERSION
    assert I.input_init("missing.txt")[0] == I.ERR_FILE_READ


def test_short_body_is_an_error(workdir):
    text = (workdir / "receiver.txt").read_text().replace("length=4", "length=9")
    (workdir / "short.txt").write_text(text)
    assert I.input_init("short.txt")[0] == I.ERR_CONFIG_DATA_READ_ELEMENT


def test_cartesian_queries():
    inp = I.Input(length=3, type=np.array([0, 1, 0], np.int32),
                  x=np.array([0.0, 100.0, 250.0], np.float32),
                  y=np.array([50.0, 0.0, 10.0], np.float32),
                  z=np.array([-100.0, -999.0, 0.0], np.float32))
    x, y, z = I.cartesian_queries(inp, 50.0, 64)
    assert np.array_equal(x, [-25.0, 75.0, 225.0])
    assert np.array_equal(y, inp.y)
    assert np.array_equal(z, [50.0 * 63 - 100.0, 50.0 * 63, 50.0 * 63])


# ---- buffer_t (test_buffer.cu) ----------------------------------------------------------

def buf(num_elements, num_device, num_host, stride):
    return I.StepBuffer(num_elements, num_device, num_host, stride, device="cpu")


def test_buffer_predicates_known_answers():
    s = 2
    b = buf(10, 2, 2, s)
    assert [b.is_device_ready(t) for t in (0, 1, s - 1, s, 2 * s)] == [1, 0, 0, 1, 1]
    b = buf(10, 1, 2, s)
    assert [b.is_device_full(t) for t in (0, 1, s, 2 * s)] == [1, 0, 1, 1]
    assert [b.is_device_empty(t) for t in (0, 1, s, 2 * s)] == [1, 0, 1, 1]
    b = buf(10, 4, 2, s)
    assert [b.is_device_full(t) for t in (0, 1, s, 3 * s, 7 * s)] == [0, 0, 0, 1, 1]
    assert [b.is_device_empty(t) for t in (0, 1, s, 4 * s, 8 * s)] == [1, 0, 0, 1, 1]
    b = buf(10, 2, 1, s)
    assert [b.is_host_full(t) for t in (0, 1, s, 3 * s, 3 * s + 1)] == [0, 0, 1, 1, 0]
    assert [b.is_host_empty(t) for t in (0, 1, s, 2 * s, 2 * s + 1)] == [1, 0, 0, 1, 0]
    b = buf(10, 2, 2, s)
    assert [b.is_host_full(t) for t in (0, 1, s, 3 * s, 3 * s + 1)] == [0, 0, 0, 1, 0]
    assert [b.is_host_empty(t) for t in (0, 1, s, 4 * s, 4 * s + 1, 8 * s, 8 * s + 1)] == \
        [1, 0, 0, 1, 0, 1, 0]


def test_buffer_offsets_known_answers():
    s, n = 2, 10
    b = buf(n, 2, 2, s)
    assert [b.device_offset(t) for t in (0, s - 1, s, s + 1)] == [0, 0, n, n]
    b = buf(n, 2, 3, s)
    assert [b.host_offset(t) for t in (0, s * 2, 2 * s * 2, 2 * s * 2 + 1)] == \
        [0, 2 * n, 4 * n, 4 * n]


def test_buffer_copy_to_host_known_answer():
    # test_buffer.cu:226-262: two fills of the host buffer, the second kept
    stride, n, nd, nh = 4, 2, 2, 3
    b = buf(n, nd, nh, stride)
    for step in range(1, 50):
        if b.is_device_ready(step):
            b.device_slot(step)[:] = float(step)
        if b.is_device_full(step):
            b.copy_to_host(step)
    assert np.array_equal(b.h_buffer, [24, 24, 28, 28, 32, 32, 36, 36, 40, 40, 44, 44])


def test_receiver_file_blocks_follow_the_indexed_view(tmp_path):
    """mpi_io_idx_write's file view (src/mpi/io.c:55-100): receiver r's block
    of write w lands at element (r * num_writes + w) * num_elements.  First
    the known answer of the reference's test_mpi_indexed_write
    (tests/mpi/test_io.c:186-240: 4 blocks of 2 elements, 2 writes of the
    same data, block i holding global_data[i]); then distinct values per
    receiver, write and element, which fix the order of all three."""
    from awp_amd.inputs import write_receiver_block
    name = str(tmp_path / "temp.bin")
    global_data = np.arange(1, 5, dtype=np.float32)
    data = np.repeat(global_data[:, None], 2, axis=1)
    for w in range(2):
        write_receiver_block(name, data, w, 2)
    buf = np.fromfile(name, dtype=np.float32)
    ans = np.zeros(16, dtype=np.float32)
    for i in range(4):
        for j in range(2):
            for k in range(2):
                ans[k + j * 2 + i * 2 * 2] = global_data[i]
    assert np.array_equal(buf, ans)
    name = str(tmp_path / "r.bin")
    nq, ne, nw = 3, 5, 3
    blocks = [np.arange(nq * ne, dtype=np.float32).reshape(nq, ne) + 100 * w for w in range(nw)]
    for w in range(nw):
        write_receiver_block(name, blocks[w], w, nw)
    got = np.fromfile(name, dtype=np.float32).reshape(nq, nw, ne)
    for r in range(nq):
        for w in range(nw):
            assert np.array_equal(got[r, w], blocks[w][r])
    # one write per file: the transposed host buffer as it is
    name = str(tmp_path / "one.bin")
    write_receiver_block(name, blocks[0], 0, 1)
    assert np.array_equal(np.fromfile(name, dtype=np.float32), blocks[0].ravel())
