"""Python mirror of the MI355X exact negacyclic product (C-ABI: include/polymul.h).

The reference exposes one C++ entry-point triple (fherma.h:46-48):
fherma_init(Point) -> state, fherma_run(state, Inputs) -> Outputs,
fherma_free(state).  `Solver` mirrors it one to one (init / run / free with
the same argument meaning and the same failure behaviour: a RuntimeError
naming the operation), on top of `Context`, which binds the C-ABI.

The product path is the HIP library `libpolymul.so` built next to this file;
nothing here falls back to a CPU implementation -- if the library is absent,
loading fails loudly.
"""
import ctypes
import os

import numpy as np

from . import pointdir  # noqa: F401  (the solution harness's point directories)
from . import workload  # noqa: F401  (seeded synthetic inputs)

LIMBS = 28
WIDTH = 868
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PM_OK, PM_EINVAL, PM_EUNSUPPORTED, PM_EDEVICE, PM_ENOMEM, PM_EBUDGET, PM_ETIMEOUT = range(7)
MAX_IO_CHUNKS = 16  # input/output chunks of the host-buffer path at most


class PolymulError(RuntimeError):
    def __init__(self, operation, status, reason):
        super().__init__(f"{operation} failed (status {status}): {reason}")
        self.status = status


class _Info(ctypes.Structure):
    _fields_ = [("max_n", ctypes.c_uint32), ("k", ctypes.c_uint32), ("product_bits", ctypes.c_uint32),
                ("device", ctypes.c_uint32), ("table_bytes", ctypes.c_uint64),
                ("workspace_bytes", ctypes.c_uint64)]


def library_path():
    return os.path.join(_HERE, "libpolymul.so")


def library():
    """Load libpolymul.so (raises OSError when it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with __graft_entry__.build() or make -C {_HERE}")
    # PyTorch-ROCm ships its own libamdhip64 and asks for it by a name that
    # the system runtime does not answer to; loaded in that order the process
    # ends up with two HIP runtimes and torch then finds no GPU.  Importing
    # torch first lets this library bind to torch's runtime (same soname).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    u32 = ctypes.c_uint32
    lib.pm_create.argtypes = [u32, u32, u32, P, ctypes.c_int, ctypes.POINTER(P)]
    lib.pm_destroy.argtypes = [P]
    lib.pm_destroy.restype = None
    lib.pm_last_error.argtypes = [P]
    lib.pm_last_error.restype = ctypes.c_char_p
    lib.pm_plan_rns.argtypes = [u32, P, ctypes.POINTER(u32), P, P, ctypes.POINTER(u32)]
    lib.pm_plan_quartic.argtypes = [u32, P, ctypes.POINTER(u32), P, P, P, ctypes.POINTER(u32)]
    lib.pm_get_info.argtypes = [P, ctypes.POINTER(_Info)]
    lib.pm_get_prime.argtypes = [P, u32, ctypes.POINTER(u32), ctypes.POINTER(u32)]
    lib.pm_multiply.argtypes = [P, u32, P, P, P]
    lib.pm_multiply_submit.argtypes = [P, u32, P, P]
    lib.pm_multiply_collect.argtypes = [P, P]
    lib.pm_set_io_chunks.argtypes = [P, u32, u32]
    lib.pm_set_io_wait_limit.argtypes = [P, u32]
    lib.pm_inject_io_stall.argtypes = [P, ctypes.c_int32]
    lib.pm_place_near_device.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.pm_multiply_device.argtypes = [P, u32, P, P, P, P]
    lib.pm_multiply_device_noalloc.argtypes = [P, u32, P, P, P, P]
    lib.pm_synchronize.argtypes = [P]
    lib.pm_set_backend.argtypes = [P, u32]
    lib.pm_get_backend.argtypes = [P, ctypes.POINTER(u32), ctypes.POINTER(u32)]
    lib.pm_set_outer.argtypes = [P, u32, u32]
    lib.pm_get_outer.argtypes = [P, u32, ctypes.POINTER(u32)]
    lib.pm_set_batch_outer.argtypes = [P, u32, u32]
    lib.pm_get_batch_outer.argtypes = [P, u32, ctypes.POINTER(u32)]
    lib.pm_set_batch_tiles_per_block.argtypes = [P, u32, u32]
    lib.pm_get_batch_tiles_per_block.argtypes = [P, u32, ctypes.POINTER(u32)]
    lib.pm_set_batch_residues_per_thread.argtypes = [P, u32, u32]
    lib.pm_get_batch_residues_per_thread.argtypes = [P, u32, ctypes.POINTER(u32)]
    lib.pm_set_batch_crt_waves.argtypes = [P, u32]
    lib.pm_set_batch_parts.argtypes = [P, u32]
    lib.pm_get_batch_parts.argtypes = [P, ctypes.POINTER(u32)]
    lib.pm_get_batch_crt_waves.argtypes = [P, ctypes.POINTER(u32)]
    lib.pm_set_conversion_slices.argtypes = [P, u32, u32]
    lib.pm_get_conversion_slices.argtypes = [P, u32, ctypes.POINTER(u32)]
    lib.pm_stage_convert.argtypes = [P, u32, P, P, P, P, P]
    lib.pm_stage_convert_scalar.argtypes = [P, u32, P, P, P, P, P]
    lib.pm_stage_ntt.argtypes = [P, u32, u32, P, P, P]
    lib.pm_stage_crt.argtypes = [P, u32, P, P, P]
    lib.pm_stage_crt_scalar.argtypes = [P, u32, P, P, P]
    lib.pm_batch_create.argtypes = [P, u32, P, P, P, P, u32, ctypes.POINTER(P)]
    lib.pm_batch_destroy.argtypes = [P]
    lib.pm_batch_destroy.restype = None
    lib.pm_batch_io.argtypes = [P, u32, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P)]
    lib.pm_batch_upload.argtypes = [P, u32, P, P]
    lib.pm_batch_download.argtypes = [P, u32, P]
    lib.pm_batch_run.argtypes = [P, P]
    lib.pm_batch_wait.argtypes = [P, P]
    lib.pm_batch_phase_times.argtypes = [P, P, u32, P, P, ctypes.POINTER(u32)]
    lib.pm_plan_worklist.argtypes = [u32, u32, P, ctypes.POINTER(u32), P, ctypes.POINTER(u32), P]
    lib.pm_batch_info.argtypes = [P, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_uint64)]
    lib.pm_server_create.argtypes = [P, ctypes.c_uint64, u32, u32, ctypes.POINTER(P)]
    lib.pm_server_submit.argtypes = [P, u32, P, P, P, ctypes.POINTER(ctypes.c_uint64)]
    lib.pm_server_wait.argtypes = [P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    lib.pm_server_stats.argtypes = [P] + [ctypes.POINTER(ctypes.c_uint64)] * 4
    lib.pm_server_destroy.argtypes = [P]
    lib.pm_server_destroy.restype = None
    for name in EXPORTED:
        getattr(lib, name)  # every declared symbol must resolve
    _LIB = lib
    return lib


# every extern "C" entry point declared in include/polymul.h
EXPORTED = ["pm_plan_rns", "pm_plan_quartic", "pm_create", "pm_destroy", "pm_last_error", "pm_get_info", "pm_get_prime", "pm_multiply",
            "pm_multiply_submit", "pm_multiply_collect", "pm_set_io_chunks", "pm_set_io_wait_limit",
            "pm_inject_io_stall", "pm_place_near_device", "pm_multiply_device", "pm_multiply_device_noalloc",
            "pm_synchronize", "pm_set_backend", "pm_get_backend", "pm_set_outer", "pm_get_outer", "pm_set_batch_outer", "pm_get_batch_outer",
            "pm_set_batch_tiles_per_block", "pm_get_batch_tiles_per_block", "pm_set_batch_residues_per_thread",
            "pm_get_batch_residues_per_thread", "pm_set_batch_crt_waves",
            "pm_get_batch_crt_waves", "pm_set_batch_parts", "pm_get_batch_parts", "pm_set_conversion_slices", "pm_get_conversion_slices",
            "pm_stage_convert",
            "pm_stage_convert_scalar", "pm_stage_ntt", "pm_stage_crt", "pm_stage_crt_scalar",
            "pm_batch_create", "pm_batch_destroy", "pm_batch_io", "pm_batch_upload", "pm_batch_download",
            "pm_batch_run", "pm_batch_wait", "pm_batch_info", "pm_plan_worklist", "pm_batch_phase_times",
            "pm_server_create", "pm_server_submit", "pm_server_wait", "pm_server_stats", "pm_server_destroy"]


def _ptr(x):
    """Address of a numpy array or a torch tensor (device or host)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.ctypes.data


def place_near_device(device=0):
    """Restricts the calling thread to the CPUs of the device's NUMA node
    (pm_place_near_device); returns (status, node, cpus the thread may use)."""
    lib = library()
    node, cpus = ctypes.c_int(-1), ctypes.c_int(0)
    status = lib.pm_place_near_device(int(device), ctypes.byref(node), ctypes.byref(cpus))
    return status, node.value, cpus.value


def plan_rns(q, max_n):
    """Host-only RNS plan: (k, primes, psi_max, product_bits); needs no GPU."""
    lib = library()
    q = np.ascontiguousarray(np.asarray(q, dtype=np.uint32).reshape(LIMBS))
    k, bits = ctypes.c_uint32(), ctypes.c_uint32()
    primes = np.zeros(64, dtype=np.uint32)
    psi = np.zeros(64, dtype=np.uint32)
    status = lib.pm_plan_rns(max_n, q.ctypes.data, ctypes.byref(k), primes.ctypes.data, psi.ctypes.data,
                             ctypes.byref(bits))
    if status != PM_OK:
        raise PolymulError("pm_plan_rns", status, lib.pm_last_error(None).decode())
    return k.value, primes[:k.value].copy(), psi[:k.value].copy(), bits.value


def plan_quartic(q, max_n, iotas=False):
    """Host-only plan of the quartic backend (q = 2^868 - delta, delta < 2^28):
    (k, primes, fourth roots of delta, product_bits), and with iotas=True also
    the fourth roots of unity of the channels; needs no GPU."""
    lib = library()
    q = np.ascontiguousarray(np.asarray(q, dtype=np.uint32).reshape(LIMBS))
    k, bits = ctypes.c_uint32(), ctypes.c_uint32()
    primes = np.zeros(32, dtype=np.uint32)
    roots = np.zeros(32, dtype=np.uint32)
    io = np.zeros(32, dtype=np.uint32)
    status = lib.pm_plan_quartic(max_n, q.ctypes.data, ctypes.byref(k), primes.ctypes.data, roots.ctypes.data,
                                 io.ctypes.data, ctypes.byref(bits))
    if status != PM_OK:
        raise PolymulError("pm_plan_quartic", status, lib.pm_last_error(None).decode())
    out = (k.value, primes[:k.value].copy(), roots[:k.value].copy(), bits.value)
    return out + (io[:k.value].copy(),) if iotas else out


def plan_worklist(k, ns):
    """Host-only work list of a batch: (items [n,4] uint32, units [m,2] uint32)."""
    lib = library()
    ns_arr = np.ascontiguousarray(ns, dtype=np.uint32)
    ni, nu = ctypes.c_uint32(0), ctypes.c_uint32(0)
    st = lib.pm_plan_worklist(k, len(ns_arr), ns_arr.ctypes.data, ctypes.byref(ni), None, ctypes.byref(nu), None)
    if st != PM_OK:
        raise PolymulError("pm_plan_worklist", st, "invalid sizes")
    items = np.zeros((ni.value, 4), dtype=np.uint32)
    units = np.zeros((max(nu.value, 1), 2), dtype=np.uint32)
    st = lib.pm_plan_worklist(k, len(ns_arr), ns_arr.ctypes.data, ctypes.byref(ni), items.ctypes.data,
                              ctypes.byref(nu), units.ctypes.data)
    if st != PM_OK:
        raise PolymulError("pm_plan_worklist", st, "invalid sizes")
    return items, units[:nu.value]


class Context:
    """An RNS product context for one modulus q (pm_create / pm_destroy)."""

    def __init__(self, q, max_n, device=0):
        lib = library()
        q = np.ascontiguousarray(np.asarray(q, dtype=np.uint32).reshape(LIMBS))
        handle = ctypes.c_void_p()
        status = lib.pm_create(max_n, WIDTH, LIMBS, q.ctypes.data, device, ctypes.byref(handle))
        if status != PM_OK:
            raise PolymulError("pm_create", status, lib.pm_last_error(None).decode())
        self._lib, self._h, self.q = lib, handle, q.copy()

    def _check(self, operation, status):
        if status != PM_OK:
            raise PolymulError(operation, status, self._lib.pm_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def info(self):
        out = _Info()
        self._check("pm_get_info", self._lib.pm_get_info(self._h, ctypes.byref(out)))
        return {f: getattr(out, f) for f, _ in _Info._fields_}

    def primes(self):
        """[(p_i, psi_max_i)] of the RNS."""
        out = []
        for i in range(self.info()["k"]):
            p, psi = ctypes.c_uint32(), ctypes.c_uint32()
            self._check("pm_get_prime", self._lib.pm_get_prime(self._h, i, ctypes.byref(p), ctypes.byref(psi)))
            out.append((p.value, psi.value))
        return out

    def multiply(self, a, b):
        """Host arrays [n, 28] uint32 -> c [n, 28] uint32 (pm_multiply)."""
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        if a.shape != b.shape or a.ndim != 2 or a.shape[1] != LIMBS:
            raise ValueError("a and b must both be [n, 28] uint32")
        c = np.empty_like(a)
        self._check("pm_multiply", self._lib.pm_multiply(self._h, a.shape[0], a.ctypes.data, b.ctypes.data,
                                                          c.ctypes.data))
        return c

    def submit(self, a, b):
        """First half of `multiply` (pm_multiply_submit): a and b may be
        reused once this returns."""
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        if a.shape != b.shape or a.ndim != 2 or a.shape[1] != LIMBS:
            raise ValueError("a and b must both be [n, 28] uint32")
        self._check("pm_multiply_submit", self._lib.pm_multiply_submit(self._h, a.shape[0], a.ctypes.data,
                                                                        b.ctypes.data))
        self._pending_n = a.shape[0]

    def collect(self, out=None):
        """Second half (pm_multiply_collect): the product of the last submit."""
        n = getattr(self, "_pending_n", None)
        if n is None:
            raise PolymulError("pm_multiply_collect", 1, "no submitted product to collect")
        c = np.empty((n, LIMBS), dtype=np.uint32) if out is None else out
        if c.shape != (n, LIMBS) or c.dtype != np.uint32 or not c.flags.c_contiguous:
            raise ValueError("out must be a contiguous [n, 28] uint32 array")
        self._pending_n = None
        self._check("pm_multiply_collect", self._lib.pm_multiply_collect(self._h, c.ctypes.data))
        return c

    def set_io_wait_limit(self, microseconds):
        """Bound of one device wait for a host-buffer input chunk (0 = 5 s)."""
        self._check("pm_set_io_wait_limit", self._lib.pm_set_io_wait_limit(self._h, int(microseconds)))

    def inject_io_stall(self, chunk):
        """Next submit never raises input chunk `chunk`'s ready flag (failure injection; -1 cancels)."""
        self._check("pm_inject_io_stall", self._lib.pm_inject_io_stall(self._h, int(chunk)))

    def set_io_chunks(self, n, chunks):
        """Chunks of the host-buffer path for products of size n (0 = default)."""
        self._check("pm_set_io_chunks", self._lib.pm_set_io_chunks(self._h, int(n), int(chunks)))

    def multiply_device(self, n, a, b, c, stream=None):
        self._check("pm_multiply_device",
                    self._lib.pm_multiply_device(self._h, n, _ptr(a), _ptr(b), _ptr(c), stream))

    def multiply_device_noalloc(self, n, a, b, c, stream=None):
        """pm_multiply_device with the residues in the library's static
        scratch: no allocation; calls overlapping on one GPU race."""
        self._check("pm_multiply_device_noalloc",
                    self._lib.pm_multiply_device_noalloc(self._h, n, _ptr(a), _ptr(b), _ptr(c), stream))

    BACKENDS = {"general": 0, "quartic": 1}

    def set_backend(self, name):
        """'general' (any odd q) or 'quartic' (q = 2^868 - delta, delta < 2^28)
        for every later product, batch and server of this context."""
        self._check("pm_set_backend", self._lib.pm_set_backend(self._h, self.BACKENDS[name]))

    def backend(self):
        """(name, residue rows per polynomial) of the context's backend."""
        b, rows = ctypes.c_uint32(), ctypes.c_uint32()
        self._check("pm_get_backend", self._lib.pm_get_backend(self._h, ctypes.byref(b), ctypes.byref(rows)))
        return {v: k for k, v in self.BACKENDS.items()}[b.value], rows.value

    def synchronize(self):
        self._check("pm_synchronize", self._lib.pm_synchronize(self._h))

    def set_outer(self, n, outer):
        """Plain-path schedule: outer radix-2 passes for products of size n."""
        self._check("pm_set_outer", self._lib.pm_set_outer(self._h, int(n), int(outer)))

    def get_outer(self, n):
        v = ctypes.c_uint32()
        self._check("pm_get_outer", self._lib.pm_get_outer(self._h, int(n), ctypes.byref(v)))
        return v.value

    def set_batch_outer(self, n, outer):
        """Batch schedule: outer radix-2 passes for products of size n in
        batches and served launches created afterwards, on the context's
        current backend (each backend keeps its own)."""
        self._check("pm_set_batch_outer", self._lib.pm_set_batch_outer(self._h, int(n), int(outer)))

    def get_batch_outer(self, n):
        v = ctypes.c_uint32()
        self._check("pm_get_batch_outer", self._lib.pm_get_batch_outer(self._h, int(n), ctypes.byref(v)))
        return v.value

    def set_batch_residues_per_thread(self, tile_n, residues):
        """Batch schedule: residues each thread of a tile of size tile_n holds
        (32, the default; or 16 for tiles of 2^6 .. 2^14); resets the size's
        tiles per block to that choice's default."""
        self._check("pm_set_batch_residues_per_thread",
                    self._lib.pm_set_batch_residues_per_thread(self._h, int(tile_n), int(residues)))

    def get_batch_residues_per_thread(self, tile_n):
        v = ctypes.c_uint32()
        self._check("pm_get_batch_residues_per_thread",
                    self._lib.pm_get_batch_residues_per_thread(self._h, int(tile_n), ctypes.byref(v)))
        return v.value

    def set_batch_tiles_per_block(self, tile_n, tiles):
        """Batch schedule: tiles of size tile_n per block (0 = the default:
        one tile from 2^13 up, a 1024-thread block of smaller tiles, at 32
        residues per thread) in batches and served launches created
        afterwards."""
        self._check("pm_set_batch_tiles_per_block",
                    self._lib.pm_set_batch_tiles_per_block(self._h, int(tile_n), int(tiles)))

    def get_batch_tiles_per_block(self, tile_n):
        v = ctypes.c_uint32()
        self._check("pm_get_batch_tiles_per_block",
                    self._lib.pm_get_batch_tiles_per_block(self._h, int(tile_n), ctypes.byref(v)))
        return v.value

    def set_batch_crt_waves(self, waves):
        """Batch schedule: waves per block of the MFMA CRT (2, 4 or 8; 0 =
        the default 8) in batches and served launches created afterwards."""
        self._check("pm_set_batch_crt_waves", self._lib.pm_set_batch_crt_waves(self._h, int(waves)))

    def get_batch_crt_waves(self):
        v = ctypes.c_uint32()
        self._check("pm_get_batch_crt_waves", self._lib.pm_get_batch_crt_waves(self._h, ctypes.byref(v)))
        return v.value

    def set_batch_parts(self, parts):
        """Batch schedule: staggered parts (1-4; 0 = chosen per batch: 3
        where a product runs 2^15 tiles, else 1) of batches created
        afterwards."""
        self._check("pm_set_batch_parts", self._lib.pm_set_batch_parts(self._h, int(parts)))

    def get_batch_parts(self):
        v = ctypes.c_uint32()
        self._check("pm_get_batch_parts", self._lib.pm_get_batch_parts(self._h, ctypes.byref(v)))
        return v.value

    def set_conversion_slices(self, n, slices):
        """Plain-path schedule: slices of the primes the conversion of a
        product of size n spreads over (1, 2 or 4; 0 = chosen at launch)."""
        self._check("pm_set_conversion_slices", self._lib.pm_set_conversion_slices(self._h, int(n), int(slices)))

    def get_conversion_slices(self, n):
        v = ctypes.c_uint32()
        self._check("pm_get_conversion_slices",
                    self._lib.pm_get_conversion_slices(self._h, int(n), ctypes.byref(v)))
        return v.value

    def schedule(self):
        """{n: outer} for every power of two up to max_n."""
        return {1 << lg: self.get_outer(1 << lg) for lg in range(1, self.info()["max_n"].bit_length())}

    def load_schedule(self, path):
        """Apply a schedule written by `autotune_plain` / `autotune_batch` /
        `autotune_backend` (sizes this context cannot hold, and a quartic
        backend its modulus does not allow, are skipped); returns {"outer":
        {n: u}, "batch_outer": {n: u}, "batch_residues_per_thread": {tile_n:
        r}, "batch_tiles_per_block": {tile_n: tiles}, "conversion_slices": {n:
        s}} and "backend", "batch_crt_waves" and "batch_parts" when they were
        set, as applied (residues per thread before tiles per block, which
        they reset)."""
        import json
        with open(path) as f:
            doc = json.load(f)
        applied = {"outer": {}, "batch_outer": {}, "batch_residues_per_thread": {}, "batch_tiles_per_block": {},
                   "conversion_slices": {}}
        if "backend" in doc:
            try:
                self.set_backend(doc["backend"])
                applied["backend"] = doc["backend"]
            except PolymulError as err:
                if err.status != PM_EUNSUPPORTED:
                    raise
        if "batch_crt_waves" in doc:
            self.set_batch_crt_waves(doc["batch_crt_waves"])
            applied["batch_crt_waves"] = int(doc["batch_crt_waves"])
        if "batch_parts" in doc:
            self.set_batch_parts(doc["batch_parts"])
            applied["batch_parts"] = int(doc["batch_parts"])
        for key, setter in (("outer", self.set_outer), ("batch_outer", self.set_batch_outer),
                            ("batch_residues_per_thread", self.set_batch_residues_per_thread),
                            ("batch_tiles_per_block", self.set_batch_tiles_per_block),
                            ("conversion_slices", self.set_conversion_slices)):
            for n_str, outer in doc.get(key, {}).items():
                n = int(n_str)
                if n <= self.info()["max_n"]:
                    setter(n, outer)
                    applied[key][n] = outer
        return applied

    def stage_convert(self, n, a, b, ra, rb, stream=None):
        self._check("pm_stage_convert",
                    self._lib.pm_stage_convert(self._h, n, _ptr(a), _ptr(b), _ptr(ra), _ptr(rb), stream))

    def stage_convert_scalar(self, n, a, b, ra, rb, stream=None):
        self._check("pm_stage_convert_scalar",
                    self._lib.pm_stage_convert_scalar(self._h, n, _ptr(a), _ptr(b), _ptr(ra), _ptr(rb), stream))

    def stage_ntt(self, n, ra, rb, outer=0, stream=None):
        self._check("pm_stage_ntt", self._lib.pm_stage_ntt(self._h, n, outer, _ptr(ra), _ptr(rb), stream))

    def stage_crt(self, n, t, c, stream=None):
        self._check("pm_stage_crt", self._lib.pm_stage_crt(self._h, n, _ptr(t), _ptr(c), stream))

    def stage_crt_scalar(self, n, t, c, stream=None):
        self._check("pm_stage_crt_scalar", self._lib.pm_stage_crt_scalar(self._h, n, _ptr(t), _ptr(c), stream))


class Batch:
    """Products of uneven N in one captured hipGraph (one launch per phase)."""

    PERSISTENT, PHASED, SCALAR_CONVERSION, SCALAR_CRT, SEPARATE_OUTER = 1, 2, 4, 8, 16

    def __init__(self, ctx, ns, d_a=None, d_b=None, d_c=None, mode="phased"):
        self._ctx, self._lib = ctx, ctx._lib
        self.ns = [int(n) for n in ns]
        count = len(self.ns)
        arr = (ctypes.c_uint32 * count)(*self.ns)
        ptrs = [None, None, None]
        if d_a is not None:
            ptrs = [(ctypes.c_void_p * count)(*[_ptr(x) for x in v]) for v in (d_a, d_b, d_c)]
        handle = ctypes.c_void_p()
        flags = {"phased": self.PHASED, "persistent": self.PERSISTENT,
                 "phased-separate-outer": self.PHASED | self.SEPARATE_OUTER,
                 "phased-scalar-conversion": self.PHASED | self.SCALAR_CONVERSION,
                 "phased-scalar": self.PHASED | self.SCALAR_CONVERSION | self.SCALAR_CRT}[mode]
        ctx._check("pm_batch_create", self._lib.pm_batch_create(ctx.handle, count, arr, ptrs[0], ptrs[1], ptrs[2],
                                                                 flags, ctypes.byref(handle)))
        self.mode = mode
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, i, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        assert a.shape == (self.ns[i], LIMBS) and b.shape == a.shape
        self._ctx._check("pm_batch_upload", self._lib.pm_batch_upload(self._h, i, a.ctypes.data, b.ctypes.data))

    def download(self, i):
        c = np.empty((self.ns[i], LIMBS), dtype=np.uint32)
        self._ctx._check("pm_batch_download", self._lib.pm_batch_download(self._h, i, c.ctypes.data))
        return c

    def run(self, stream=None):
        self._ctx._check("pm_batch_run", self._lib.pm_batch_run(self._h, stream))

    def wait(self, stream=None):
        self._ctx._check("pm_batch_wait", self._lib.pm_batch_wait(self._h, stream))

    PHASE_NAMES = {0: "conversion", 1: "outer_forward", 3: "outer_inverse", 4: "crt", 5: "conversion", 6: "crt"}

    def phase_times(self, stream=None):
        """One direct (ungraphed) run with events between phases: [(name, ms)]."""
        kinds = np.zeros(32, dtype=np.uint32)
        ms = np.zeros(32, dtype=np.float64)
        count = ctypes.c_uint32()
        self._ctx._check("pm_batch_phase_times", self._lib.pm_batch_phase_times(
            self._h, stream, 32, kinds.ctypes.data, ms.ctypes.data, ctypes.byref(count)))
        out = []
        for kd, t in zip(kinds[:count.value], ms[:count.value]):
            kd = int(kd)
            name = f"tile_2^{(kd >> 8) & 0xFF}" if (kd & 0xFF) == 2 else self.PHASE_NAMES[kd & 0xFF]
            out.append((name, float(t)))
        return out

    def info(self):
        items, parts, nbytes = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        self._ctx._check("pm_batch_info", self._lib.pm_batch_info(self._h, ctypes.byref(items), ctypes.byref(parts),
                                                                   ctypes.byref(nbytes)))
        return {"items": items.value, "parts": parts.value, "device_bytes": nbytes.value}


class Server:
    """Continuous-batching request loop (pm_server_*)."""

    def __init__(self, ctx, hbm_budget_bytes, max_batch=64, max_inflight=2):
        self._ctx, self._lib = ctx, ctx._lib
        handle = ctypes.c_void_p()
        ctx._check("pm_server_create", self._lib.pm_server_create(ctx.handle, int(hbm_budget_bytes), max_batch,
                                                                   max_inflight, ctypes.byref(handle)))
        self._h = handle
        self._keep = {}

    def submit(self, a, b, out=None):
        """Queue a*b; returns a request id.  The arrays stay referenced until
        wait().  `out` (optional, same shape as a, uint32, C-contiguous) receives
        the product; page-locked memory there lets the result come back
        asynchronously behind the kernels.  Page-locked a and b are read by a
        kernel of the launch (no runtime copy)."""
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        if out is None:
            c = np.empty_like(a)
        else:
            c = out
            if c.dtype != np.uint32 or c.shape != a.shape or not c.flags.c_contiguous or not c.flags.writeable:
                raise ValueError("out must be a writeable C-contiguous uint32 array shaped like a")
        rid = ctypes.c_uint64()
        self._ctx._check("pm_server_submit", self._lib.pm_server_submit(self._h, a.shape[0], a.ctypes.data,
                                                                         b.ctypes.data, c.ctypes.data,
                                                                         ctypes.byref(rid)))
        self._keep[rid.value] = (a, b, c)
        return rid.value

    def wait(self, rid):
        """Block until request rid is done: (c, latency_seconds)."""
        lat = ctypes.c_double()
        st = self._lib.pm_server_wait(self._h, rid, ctypes.byref(lat))
        a, b, c = self._keep.pop(rid)
        self._ctx._check("pm_server_wait", st)
        return c, lat.value

    def stats(self):
        vals = [ctypes.c_uint64() for _ in range(4)]
        self._ctx._check("pm_server_stats", self._lib.pm_server_stats(self._h, *[ctypes.byref(v) for v in vals]))
        return dict(zip(["completed", "launches", "max_concurrent", "peak_bytes"], [v.value for v in vals]))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_plain(ctx, ns, reps=20, path=None, stream=None):
    """Per-N schedule of the plain path, timed on the GPU one product at a
    time: first the outer radix-2 passes (0, 1, 2, where the tile stays
    within 2..2^15), then, with the best of those, the slices of the primes
    the conversion spreads over (1, 2, 4).  Keeps the fastest in the context
    and returns {n: {"outer": u, "ms": {u: median ms}, "slices": s,
    "slices_ms": {s: median ms}}}; with `path` the choice is also written as
    JSON ({"device": ..., "outer": {n: u}, "conversion_slices": {n: s}}) for
    `load_schedule`."""
    import torch
    from . import workload as wl
    s = stream if stream is not None else torch.cuda.Stream()

    def timed(n, a, b, c):
        with torch.cuda.stream(s):
            for _ in range(3):
                ctx.multiply_device(n, a, b, c, s.cuda_stream)
            ms = []
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(s)
                ctx.multiply_device(n, a, b, c, s.cuda_stream)
                e1.record(s)
                e1.synchronize()
                ms.append(e0.elapsed_time(e1))
        return float(np.median(ms))

    result = {}
    for n in sorted(set(int(x) for x in ns)):
        logn = n.bit_length() - 1
        q = ctx.q
        a = torch.from_numpy(wl.poly(n, 0, n, q).view(np.int32)).cuda()
        b = torch.from_numpy(wl.poly(n, 1, n, q).view(np.int32)).cuda()
        c = torch.empty_like(a)
        times = {}
        for outer in (0, 1, 2):
            if outer > logn or logn - outer > 15 or logn - outer < 1:
                continue
            ctx.set_outer(n, outer)
            times[outer] = timed(n, a, b, c)
        best = min(times, key=times.get)
        ctx.set_outer(n, best)
        slice_ms = {}
        for sl in (1, 2, 4):
            ctx.set_conversion_slices(n, sl)
            slice_ms[sl] = timed(n, a, b, c)
        best_sl = min(slice_ms, key=slice_ms.get)
        ctx.set_conversion_slices(n, best_sl)
        result[n] = {"outer": best, "ms": times, "slices": best_sl, "slices_ms": slice_ms}
    if path is not None:
        _save_schedule(path, "outer", {n: r["outer"] for n, r in result.items()})
        _save_schedule(path, "conversion_slices", {n: r["slices"] for n, r in result.items()})
    return result


def _save_schedule(path, key, table):
    """Write table {n: outer} under `key` of the schedule file, keeping what
    the file already holds for the other key."""
    import json
    import torch
    doc = {}
    if os.path.exists(path):
        with open(path) as f:
            doc = json.load(f)
    doc["device"] = torch.cuda.get_device_name(0)
    doc.setdefault(key, {}).update({str(n): int(u) for n, u in table.items()})
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)


def _time_batch(ctx, ns, inputs, reps, stream):
    """Median ms of a graph launch of a batch of these sizes (schedule as set)."""
    import torch
    batch = Batch(ctx, ns)
    for i, (a, b) in enumerate(inputs):
        batch.upload(i, a, b)
    with torch.cuda.stream(stream):
        for _ in range(2):
            batch.run(stream.cuda_stream)
        batch.wait(stream.cuda_stream)
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            batch.run(stream.cuda_stream)
            e1.record(stream)
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        batch.wait(stream.cuda_stream)
    batch.close()
    return float(np.median(ms))


def tiles_per_block_choices(tile_n, residues=32):
    """Valid tiles per block of a tile size: powers of two with
    64 r <= tiles x tile_n <= 1024 r (whole waves of r residues per thread,
    at most 1024 threads)."""
    out, t = [], 1
    while t * tile_n <= 1024 * residues:
        if t * tile_n >= 64 * residues:
            out.append(t)
        t *= 2
    return out


def residues_per_thread_choices(tile_n):
    """Residues per thread a tile size can run: 32, and 16 for 2^6 .. 2^14."""
    return [32, 16] if 64 <= tile_n <= 16384 else [32]


def autotune_batch(ctx, ns, count=16, reps=5, path=None, stream=None, coefficients=1 << 20):
    """Batch schedule, the role of the reference's compile-time tile, radix
    and primes-per-block knobs (quartic/host_setup.h:11-12,
    quartic/solve.cu:325,576), chosen on the device.  Each trial is a graph
    launch of a batch of one size n holding max(count, coefficients / n)
    products, at most 512 (a batch too small to fill the device would favour
    whatever spreads a few products widest):

    1. residues per thread (32 or 16) and tiles per block of each tile size
       2^t the sizes can use, on products of N = 2^t (no outer stages);
    2. outer stages (0..2: the tile size, and radix-2 or radix-4 outer
       stages in the conversion and the CRT) for each N, with the tiles per
       block of step 1;
    3. waves per block of the MFMA CRT (2, 4, 8): the sum over the sizes;
    4. staggered parts of the graph (1-4): the batch of `ns` as given (the
       mix the caller runs; each part holds at least 4 products), compared
       with the per-batch default (0).

    All are kept in the context (for batches and served launches created
    afterwards, on its current backend).  Returns {n: {"outer": u, "ms": {u:
    median ms per launch}}, "tiles_per_block": {tile_n: {"residues": r,
    "tiles": t, "ms": {"r/t": ms}}}, "crt_waves": {"waves": w, "ms": {w: ms}}, "parts": {"parts":
    p, "ms": {p: ms}}}; with `path` the choice is also written to the
    schedule file `load_schedule` reads ("batch_outer",
    "batch_residues_per_thread", "batch_tiles_per_block", "batch_crt_waves",
    "batch_parts")."""
    import json
    import torch
    from . import workload as wl
    s = stream if stream is not None else torch.cuda.Stream()
    ns = [int(x) for x in ns]
    sizes = sorted(set(ns))
    max_n = ctx.info()["max_n"]
    polys = {}

    def operands(n):
        if n not in polys:
            polys[n] = (wl.poly(n, 0, n, ctx.q), wl.poly(n, 1, n, ctx.q))
        return polys[n]

    def trial(n):
        m = min(512, max(count, coefficients // n))
        return _time_batch(ctx, [n] * m, [operands(n)] * m, reps, s)

    tiles = {}
    for n in sizes:
        lg = n.bit_length() - 1
        for outer in (0, 1, 2):
            if outer <= lg and 5 <= lg - outer <= 15:
                tiles.setdefault(1 << (lg - outer), None)
    tile_result = {}
    for tile_n in sorted(tiles):
        combos = [(r, t) for r in residues_per_thread_choices(tile_n) for t in tiles_per_block_choices(tile_n, r)]
        if tile_n > max_n or len(combos) < 2:
            continue
        saved = ctx.get_batch_outer(tile_n)
        ctx.set_batch_outer(tile_n, 0)
        times = {}
        for r, t in combos:
            ctx.set_batch_residues_per_thread(tile_n, r)
            ctx.set_batch_tiles_per_block(tile_n, t)
            times[(r, t)] = trial(tile_n)
        best_r, best_t = min(times, key=times.get)
        ctx.set_batch_residues_per_thread(tile_n, best_r)
        ctx.set_batch_tiles_per_block(tile_n, best_t)
        ctx.set_batch_outer(tile_n, saved)
        tile_result[tile_n] = {"residues": best_r, "tiles": best_t,
                               "ms": {f"{r}/{t}": ms for (r, t), ms in times.items()}}
    result = {}
    for n in sizes:
        logn = n.bit_length() - 1
        times = {}
        for outer in (0, 1, 2):
            if outer > logn or logn - outer > 15 or logn - outer < 5:
                continue
            ctx.set_batch_outer(n, outer)
            times[outer] = trial(n)
        best = min(times, key=times.get)
        ctx.set_batch_outer(n, best)
        result[n] = {"outer": best, "ms": times}
    wave_ms = {}
    for w in (2, 4, 8):
        ctx.set_batch_crt_waves(w)
        wave_ms[w] = sum(trial(n) for n in sizes)
    best_w = min(wave_ms, key=wave_ms.get)
    ctx.set_batch_crt_waves(best_w)
    part_ms = {}
    for p in (0, 1, 2, 3, 4):
        if p and len(ns) < 4 * p:
            continue
        ctx.set_batch_parts(p)
        part_ms[p] = _time_batch(ctx, ns, [operands(n) for n in ns], reps, s)
    best_p = min(part_ms, key=part_ms.get)
    ctx.set_batch_parts(best_p)
    if path is not None:
        _save_schedule(path, "batch_outer", {n: r["outer"] for n, r in result.items()})
        _save_schedule(path, "batch_residues_per_thread", {t: r["residues"] for t, r in tile_result.items()})
        _save_schedule(path, "batch_tiles_per_block", {t: r["tiles"] for t, r in tile_result.items()})
        with open(path) as f:
            doc = json.load(f)
        doc["batch_crt_waves"] = best_w
        doc["batch_parts"] = best_p
        with open(path, "w") as f:
            json.dump(doc, f, indent=1)
    result["tiles_per_block"] = tile_result
    result["crt_waves"] = {"waves": best_w, "ms": wave_ms}
    result["parts"] = {"parts": best_p, "ms": part_ms}
    return result


def autotune_backend(ctx, n, count=16, reps=5, path=None, stream=None):
    """The backend the batched path runs fastest for products of size n:
    times a graph launch of `count` products with each backend the context's
    q allows (the quartic one only for q = 2^868 - delta, delta < 2^28), keeps
    the faster in the context and returns {"backend": name, "ms": {name:
    median ms per launch}}; with `path` the choice is written as "backend" in
    the schedule file `load_schedule` reads."""
    import json
    import torch
    from . import workload as wl
    s = stream if stream is not None else torch.cuda.Stream()
    a, b = wl.poly(n, 0, n, ctx.q), wl.poly(n, 1, n, ctx.q)
    times = {}
    for name in ("general", "quartic"):
        try:
            ctx.set_backend(name)
        except PolymulError:
            continue  # not a special modulus
        times[name] = _time_batch(ctx, [n] * count, [(a, b)] * count, reps, s)
    best = min(times, key=times.get)
    ctx.set_backend(best)
    if path is not None:
        doc = {}
        if os.path.exists(path):
            with open(path) as f:
                doc = json.load(f)
        doc["device"] = torch.cuda.get_device_name(0)
        doc["backend"] = best
        with open(path, "w") as f:
            json.dump(doc, f, indent=1)
    return {"backend": best, "ms": times}


class Solver:
    """fherma_init / fherma_run / fherma_free (fherma.h:46-48) in Python."""

    def __init__(self, N, W, L, q):
        if N < 2 or N & (N - 1) or N > 131072 or W != WIDTH or L != LIMBS or len(q) != LIMBS:
            raise RuntimeError("coverage: power-of-two 2<=N<=131072, W=868, L=28")
        self.N = N
        self._ctx = Context(q, N)

    def run(self, a, b):
        a = np.asarray(a, dtype=np.uint32).reshape(-1, LIMBS)
        b = np.asarray(b, dtype=np.uint32).reshape(-1, LIMBS)
        if a.shape != b.shape:
            raise RuntimeError("fherma_run: a and b must both hold N x 28 words")
        try:
            return self._ctx.multiply(a, b)
        except PolymulError as e:
            raise RuntimeError(f"fherma_run: {e}") from e

    def free(self):
        self._ctx.close()
