"""The v2 receiver/source input files and the step buffers of the
Lagrange-interpolated path (reference src/readers/input.c,
src/readers/version.c, src/buffers/buffer.c, include/readers/input.h).

An input file is ASCII: a version line ("2.0.0"), `key=value` header lines
(`#` starts a comment line), the keyword `coordinates`, then `length` rows of
`type x y z` (type 0: volume coordinates, 1: surface coordinates).  Readers
return the reference's error codes (include/awp/error.h) rather than raising,
so callers can mirror its control flow; `read_input` raises on them.

Only what one rank needs is here: the MPI broadcast (input_broadcast) and the
multi-grid distribution of sources are not.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

INPUT_MAJOR, INPUT_MINOR, INPUT_PATCH = 2, 1, 0   # include/readers/input.h:8-10
INPUT_VOLUME_COORD, INPUT_SURFACE_COORD = 0, 1      # include/readers/input.h:81

# include/awp/error.h:6-35
SUCCESS = 0
ERR_FILE_READ = 101
ERR_FILE_WRITE = 102
ERR_GET_VERSION = 200
ERR_WRONG_VERSION = 201
ERR_CONFIG_DATA_READ_ELEMENT = 305
ERR_CONFIG_PARSE_ARG = 308
ERR_CONFIG_PARSE_UNKNOWN_ARG = 309
ERR_CONFIG_PARSE_NOT_DIVISIBLE = 311


def _atoi(s: str) -> int:
    """C atoi: optional sign and leading digits, 0 if none."""
    s = s.lstrip()
    n, i = 0, 0
    sign = 1
    if i < len(s) and s[i] in "+-":
        sign = -1 if s[i] == "-" else 1
        i += 1
    while i < len(s) and s[i].isdigit():
        n = n * 10 + int(s[i])
        i += 1
    return sign * n


def _atof(s: str) -> float:
    """C atof: the longest leading decimal float ("1e3" -> 1000), 0 if none."""
    s = s.lstrip()
    for end in range(len(s), 0, -1):
        try:
            return float(s[:end])
        except ValueError:
            continue
    return 0.0


def _tokens(s: str, delim: str) -> list[str]:
    """strtok: the non-empty pieces between delimiters."""
    return [t for t in s.split(delim) if t]


@dataclass
class Version:
    major: int = 0
    minor: int = 0
    patch: int = 0

    @classmethod
    def parse(cls, text: str) -> "Version":
        """version_init (version.c:9-17): the first three '.'-separated
        fields through atoi; a missing field stays 0."""
        v = cls()
        t = _tokens(text, ".")
        for name, tok in zip(("major", "minor", "patch"), t[:3]):
            setattr(v, name, _atoi(tok))
        return v

    def __str__(self) -> str:
        return f"{self.major}.{self.minor}.{self.patch}"

    def compatible(self, other: "Version") -> bool:
        """Only the major version must match (version.c:40-48)."""
        return self.major == other.major


CURRENT = Version(INPUT_MAJOR, INPUT_MINOR, INPUT_PATCH)


@dataclass
class Input:
    """input_t with input_init_default's defaults (input.c:18-33)."""
    version: Version = field(default_factory=Version)
    file: str = "default"
    length: int = 0
    gpu_buffer_size: int = 0
    cpu_buffer_size: int = 0
    steps: int = 1
    num_writes: int = 1
    stride: int = 1
    degree: int = 0
    type: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    x: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    y: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    z: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))

    def parse(self, line: str) -> int:
        """input_parse (input.c:125-165): one `key=value` header line."""
        err, variable, value = parse_arg(line)
        if err != SUCCESS:
            return err
        if variable == "file":
            self.file = value
        elif variable in ("length", "steps", "num_writes"):
            setattr(self, variable, int(_atof(value)))
        elif variable == "stride":
            self.stride = int(_atof(value))
        elif variable in ("degree", "gpu_buffer_size", "cpu_buffer_size"):
            setattr(self, variable, _atoi(value))
        else:
            return ERR_CONFIG_PARSE_UNKNOWN_ARG
        return SUCCESS

    def check_header(self) -> int:
        """input_check_header (input.c:322-333): with more than one step, the
        step count must be a multiple of gpu_buffer_size x cpu_buffer_size.
        (A zero product, a division by zero in the reference, fails too.)"""
        if self.steps > 1:
            per = self.gpu_buffer_size * self.cpu_buffer_size
            if per == 0 or self.steps % per != 0:
                return ERR_CONFIG_PARSE_NOT_DIVISIBLE
        return SUCCESS

    def equals(self, other: "Input") -> bool:
        """input_equals (input.c:196-211): the header fields only."""
        return (self.version == other.version and self.file == other.file and
                self.length == other.length and
                self.gpu_buffer_size == other.gpu_buffer_size and
                self.cpu_buffer_size == other.cpu_buffer_size and self.steps == other.steps and
                self.num_writes == other.num_writes and self.stride == other.stride and
                self.degree == other.degree)

    def write(self, filename: str) -> int:
        """input_write (input.c:43-79)."""
        err = file_writeable(filename)
        if err:
            return err
        with open(filename, "w") as fp:
            fp.write(f"{self.version}\n")
            fp.write(f"file={self.file}\n")
            for key in ("length", "gpu_buffer_size", "cpu_buffer_size", "steps", "num_writes",
                        "stride", "degree"):
                fp.write(f"{key}={getattr(self, key)}\n")
            fp.write("coordinates\n")
            for t, x, y, z in zip(self.type, self.x, self.y, self.z):
                fp.write("%d %g %g %g\n" % (int(t), float(x), float(y), float(z)))
        return SUCCESS


def parse_arg(line: str):
    """input_parse_arg (input.c:167-194): (err, variable, value) split at
    '=' as strtok does (empty pieces skipped, a third piece ignored)."""
    t = _tokens(line, "=")
    if len(t) < 2:
        return ERR_CONFIG_PARSE_ARG, "", ""
    return SUCCESS, t[0], t[1]


def file_readable(filename: str) -> int:
    return SUCCESS if os.path.isfile(filename) and os.access(filename, os.R_OK) else ERR_FILE_READ


def file_writeable(filename: str) -> int:
    """input_file_writeable (input.c:99-123): an existing file, or one that
    can be created (and is removed again)."""
    if file_readable(filename) == SUCCESS:
        return SUCCESS
    try:
        with open(filename, "w"):
            pass
    except OSError:
        return ERR_FILE_WRITE
    os.remove(filename)
    return SUCCESS


def input_init(filename: str):
    """input_init (input.c:35-41) -> (err, Input).  Reads the version line,
    the header up to `coordinates` (skipping `#` lines), checks that the data
    file named by `file` could be written and that steps divide into the
    buffers, then reads `length` rows of `type x y z`."""
    out = Input()
    try:
        fp = open(filename, "r")
    except OSError:
        return ERR_FILE_READ, out
    with fp:
        text = fp.read()
    lines = text.split("\n")
    if not text:
        return ERR_GET_VERSION, out
    out.version = Version.parse(lines[0])
    if out.version != CURRENT and not out.version.compatible(CURRENT):
        return ERR_WRONG_VERSION, out
    # header: whitespace-separated tokens (fscanf "%s"), comment lines skipped
    rest = lines[1:]
    li = 0
    err = SUCCESS
    body = None
    while li < len(rest):
        line = rest[li]
        li += 1
        if line.lstrip().startswith("#"):
            continue
        for tok in line.split():
            if tok.startswith("#"):      # the rest of the line is a comment
                break
            if tok == "coordinates":
                body = li
                break
            err = out.parse(tok)
            if err != SUCCESS:
                return err, out
        if body is not None:
            break
    err = file_writeable(out.file)
    if err:
        return err, out
    err = out.check_header()
    if err:
        return err, out
    # body: fscanf("%d %g %g %g") per row, whitespace-insensitive
    toks = " ".join(rest[body:] if body is not None else []).split()
    n = out.length
    out.type = np.zeros(n, np.int32)
    out.x, out.y, out.z = (np.zeros(n, np.float32) for _ in range(3))
    for i in range(n):
        row = toks[4 * i:4 * i + 4]
        if len(row) != 4:
            return ERR_CONFIG_DATA_READ_ELEMENT, out
        try:
            out.type[i] = int(row[0])
            out.x[i], out.y[i], out.z[i] = (np.float32(float(v)) for v in row[1:])
        except ValueError:
            return ERR_CONFIG_DATA_READ_ELEMENT, out
    return SUCCESS, out


def read_input(filename: str) -> Input:
    err, out = input_init(filename)
    if err != SUCCESS:
        raise ValueError(f"{filename}: input error {err}")
    return out


def cartesian_queries(inp: Input, h: float, nz: int):
    """Query coordinates of a single-grid, single-rank Cartesian mesh
    (source.c:182-198, 332-348): x shifted by -h/2 so x = 0 falls on a
    material point, z measured up from the bottom node level, so a volume
    coordinate z (<= 0 below the surface) maps to z + h (nz - 1) and a
    surface coordinate to the top node level h (nz - 1)."""
    h32 = np.float32(h)
    x = (inp.x + np.float32(-0.5) * h32).astype(np.float32)
    y = inp.y.astype(np.float32)
    top = np.float32(h32 * np.float32(nz - 1))
    z = np.where(inp.type == INPUT_SURFACE_COORD, top, inp.z - (-top)).astype(np.float32)
    return x, y, z


class StepBuffer:
    """buffer_t (src/buffers/buffer.c): num_device records of num_elements
    values on the device, num_host device-fulls on the host; a record every
    `stride` steps."""

    def __init__(self, num_elements: int, num_device: int, num_host: int, stride: int,
                 device: str = "cuda"):
        import torch
        self.num_elements, self.num_device = int(num_elements), int(num_device)
        self.num_host, self.stride = int(num_host), int(stride)
        self.d_buffer = torch.zeros(self.num_device * self.num_elements, dtype=torch.float32,
                                    device=device)
        self.h_buffer = np.zeros(self.num_host * self.num_device * self.num_elements, np.float32)

    def device_offset(self, step: int) -> int:
        return ((step // self.stride) % self.num_device) * self.num_elements

    def host_offset(self, step: int) -> int:
        return ((step // self.stride // self.num_device) % self.num_host) * \
            self.num_device * self.num_elements

    def device_slot(self, step: int):
        """buffer_get_device_ptr: this step's record on the device."""
        o = self.device_offset(step)
        return self.d_buffer[o:o + self.num_elements]

    def host_block(self, step: int) -> np.ndarray:
        """buffer_get_host_ptr: the host block this device-full maps to."""
        o = self.host_offset(step)
        return self.h_buffer[o:o + self.num_device * self.num_elements]

    def is_device_ready(self, step: int) -> bool:
        return step % self.stride == 0

    def is_device_full(self, step: int) -> bool:
        return step % self.stride == 0 and \
            (step // self.stride) % self.num_device == self.num_device - 1

    def is_device_empty(self, step: int) -> bool:
        return step % self.stride == 0 and (step // self.stride) % self.num_device == 0

    def is_host_full(self, step: int) -> bool:
        return self.is_device_full(step) and \
            (step // self.stride // self.num_device) % self.num_host == self.num_host - 1

    def is_host_empty(self, step: int) -> bool:
        return self.is_device_empty(step) and \
            (step // self.stride // self.num_device) % self.num_host == 0

    def copy_to_device(self, step: int) -> None:
        import torch
        if not self.is_device_empty(step):
            return
        self.d_buffer.copy_(torch.from_numpy(self.host_block(step)))

    def copy_to_host(self, step: int) -> None:
        if not self.is_device_full(step):
            return
        self.host_block(step)[:] = self.d_buffer.cpu().numpy()


def receiver_step(gpu_buffer_size: int, cpu_buffer_size: int, num_writes: int, stride: int,
                  step: int) -> int:
    """receiver_step (src/topography/receivers/receiver.c:106-113): the last
    step of the file that step's output goes to."""
    unit = gpu_buffer_size * cpu_buffer_size * num_writes * stride
    return (1 + step // unit) * unit


def receiver_filename(base: str, inp: Input, step: int, maxsteps: int) -> str:
    """receiver_filename (receiver.c:93-104): base_<file step>, zero-padded
    to ceil(log10(maxsteps)) + 1 digits."""
    import math
    filestep = receiver_step(inp.gpu_buffer_size, inp.cpu_buffer_size, inp.num_writes,
                             inp.stride, step)
    width = int(math.ceil(math.log10(float(maxsteps)))) + 1
    return f"{base}_{filestep:0{width}d}"


def write_receiver_block(name: str, data: np.ndarray, current_write: int,
                         num_writes: int) -> None:
    """One host-full of receiver samples into a receiver file, as
    mpi_io_idx_write places it (src/mpi/io.c:55-100 with the offsets of
    receiver_init_indexed, receiver.c:80-91): receiver r's block of write w
    (data[r], num_elements samples) goes to element offset
    (r * num_writes + w) * num_elements, so each receiver's num_writes
    blocks lie together.  data: (receivers, num_elements) float32, receiver
    r being the r-th receiver of the input file."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    nq, ne = data.shape
    if not 0 <= current_write < num_writes:
        raise ValueError(f"write {current_write} of {num_writes}")
    if num_writes == 1:
        with open(name, "wb") as fp:
            data.tofile(fp)
        return
    size = nq * num_writes * ne * 4
    if not os.path.exists(name) or os.path.getsize(name) < size:
        with open(name, "ab") as fp:
            fp.truncate(size)
    mm = np.memmap(name, dtype=np.float32, mode="r+", shape=(nq, num_writes, ne))
    mm[:, current_write, :] = data
    mm.flush()
    del mm


class ReceiverWriter:
    """receiver_write (receiver.c:50-76) for one rank: sample into the device
    buffer when the step is due, flush device-fulls to the host, and when the
    host buffer is full write it into the current file, transposed to
    receiver-major order; num_writes host-fulls share one file, each
    receiver's blocks together (write_receiver_block)."""

    def __init__(self, interp, inp: Input, base: str, maxsteps: int, directory: str = "."):
        self.interp, self.inp, self.base = interp, inp, base
        self.maxsteps, self.directory = int(maxsteps), directory
        self.buf = StepBuffer(interp.num_query, inp.gpu_buffer_size, inp.cpu_buffer_size,
                              inp.stride)
        self.current_write = 0
        self.files: list[str] = []

    def write(self, step: int, field) -> None:
        import torch
        b = self.buf
        if b.is_device_ready(step):
            self.interp.interp(b.device_slot(step), field)
        if b.is_device_full(step):
            torch.cuda.current_stream().synchronize()
            b.copy_to_host(step)
        if b.is_host_full(step):
            rows = b.num_host * b.num_device
            data = b.h_buffer.reshape(rows, b.num_elements).T.copy()
            name = os.path.join(self.directory, receiver_filename(self.base, self.inp, step,
                                                                  self.maxsteps))
            if self.current_write == 0 and os.path.exists(name):
                os.remove(name)     # a new file of num_writes host-fulls
            write_receiver_block(name, data, self.current_write, self.inp.num_writes)
            if name not in self.files:
                self.files.append(name)
            self.current_write += 1
            if self.current_write == self.inp.num_writes:
                self.current_write = 0
