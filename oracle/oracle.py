"""ctypes front for the test oracle (kd_oracle.c) and the compiled reference.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker; the product never imports it.

Two CPU checkers live behind this module:

* ``liboracle.so`` -- our own C restatement of src/kdbuild.c, src/kdquery.c
  and src/coordmath.c (see kd_oracle.c for the line-by-line citations).
* ``_ref/libkdtree_ref.so`` / ``_ref/libcoords_ref.so`` -- the reference's own
  C files compiled from their sources by oracle/Makefile, when present.  The
  reference ``pair_count`` exits the process for ``num_threads > N/20``
  (src/kdquery.c:176-179) and ``tree_construct`` exits for 256 or more copies
  of one point (src/kdbuild.c:501-511); the wrappers below refuse those inputs
  instead of calling it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_double, c_int, c_longlong, c_size_t, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

NODE_DTYPE = np.dtype(
    [
        ("value", "<f8", (3,)),
        ("field", "u1"),
        ("weight", "u1"),
        ("pad0", "u1", (6,)),
        ("bits", "<u4"),
        ("pad1", "<u4"),
    ]
)
assert NODE_DTYPE.itemsize == 40


class OrTree(Structure):
    _fields_ = [
        ("nodes", c_void_p),
        ("size", c_int),
        ("memsize", c_int),
        ("num_fields", c_int),
        ("unique", c_void_p),
        ("args", c_void_p * 3),
        ("perm", POINTER(c_int)),
        ("inorder", POINTER(c_int)),
    ]


class OrDatum(Structure):
    _fields_ = [("value", c_double * 3), ("field", c_int), ("weight", c_int), ("first", c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise OSError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.or_tree_build.restype = c_int
        L.or_tree_build.argtypes = [POINTER(c_double), POINTER(c_int), c_int, c_int, POINTER(OrTree)]
        L.or_tree_free.restype = None
        L.or_tree_free.argtypes = [POINTER(OrTree)]
        for nm in ("or_pair_count",):
            f = getattr(L, nm)
            f.restype = c_int
            f.argtypes = [POINTER(OrTree), POINTER(c_double), POINTER(c_int), c_int, c_int,
                          c_double, c_int, POINTER(c_longlong)]
        L.or_pair_count_v.restype = c_int
        L.or_pair_count_v.argtypes = [POINTER(OrTree), POINTER(c_double), POINTER(c_int), c_int, c_int,
                                      c_double, c_int, POINTER(c_longlong), POINTER(c_longlong)]
        L.or_pair_count_brute.restype = c_int
        L.or_pair_count_brute.argtypes = [POINTER(OrTree), POINTER(c_double), POINTER(c_int), c_int,
                                          c_int, c_double, c_int, POINTER(c_longlong)]
        for nm in ("or_radec2cart64",):
            f = getattr(L, nm)
            f.restype = None
            f.argtypes = [POINTER(c_double), POINTER(c_double), POINTER(c_double), c_int]
        L.or_radecdist2cart64.restype = None
        L.or_radecdist2cart64.argtypes = [POINTER(c_double)] * 4 + [c_int]
        L.or_radec2sph64.restype = None
        L.or_radec2sph64.argtypes = [POINTER(c_double)] * 4 + [c_int]
        L.or_tree_struct_size.restype = c_size_t
        L.or_node_struct_size.restype = c_size_t
        assert L.or_node_struct_size() == 40
        assert L.or_tree_struct_size() == ctypes.sizeof(OrTree)
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_int))


class OracleTree:
    """A tree built by the restatement of tree_construct (src/kdbuild.c:451-537).

    Attributes mirror the parity artefacts: ``nodes`` (structured, memsize
    records, slot 0 unused), ``size``, ``memsize``, ``unique`` (value/field/
    weight/first in (x,y,z) order), ``args`` (the presorted argsort arrays),
    ``perm`` (heap index -> unique index, -1 for empty slots) and ``inorder``
    (heap index -> segment median position)."""

    def __init__(self, xyz, fields=None, num_fields=1):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.points = xyz
        self.fields = None if fields is None else np.ascontiguousarray(fields, dtype=np.int32)
        self.num_fields = int(num_fields)
        self._t = OrTree()
        rc = lib().or_tree_build(_dp(xyz), _ip(self.fields), xyz.shape[0], self.num_fields,
                                 ctypes.byref(self._t))
        if rc == 1:
            raise ValueError("Invalid weights: 256 or more copies of one point")
        if rc == 2:
            raise ValueError("NaN coordinate")
        if rc == 3:
            raise ValueError("empty point set")
        self.size = self._t.size
        self.memsize = self._t.memsize
        buf = (ctypes.c_char * (40 * self.memsize)).from_address(self._t.nodes)
        self.nodes = np.frombuffer(bytes(buf), dtype=NODE_DTYPE).copy()
        ub = (OrDatum * self.size).from_address(self._t.unique)
        u = np.frombuffer(bytes(ub), dtype=np.dtype([("value", "<f8", (3,)), ("field", "<i4"),
                                                    ("weight", "<i4"), ("first", "<i4"),
                                                    ("pad", "<i4")]))
        self.unique = u.copy()
        self.args = [np.ctypeslib.as_array((c_int * self.size).from_address(self._t.args[d])).copy()
                     for d in range(3)]
        self.perm = np.ctypeslib.as_array(self._t.perm, shape=(self.memsize,)).copy()
        self.inorder = np.ctypeslib.as_array(self._t.inorder, shape=(self.memsize,)).copy()

    def pair_count(self, q, qfields=None, nq_fields=1, r=0.1, nthreads=8):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        qf = None if qfields is None else np.ascontiguousarray(qfields, dtype=np.int32)
        out = np.zeros(nq_fields * self.num_fields, dtype=np.int64)
        bad = lib().or_pair_count(ctypes.byref(self._t), _dp(q), _ip(qf), q.shape[0], int(nq_fields),
                                  float(r), int(nthreads), out.ctypes.data_as(POINTER(c_longlong)))
        if bad:
            raise IndexError("field index outside the count matrix")
        return out.reshape(nq_fields, self.num_fields)

    def traversal_visits(self, q, qfields=None, nq_fields=1, r=0.1, nthreads=8):
        """Number of node distances the reference's traversal computes for
        these query points at radius r (one per visited node, kdquery.c:84)."""
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        qf = None if qfields is None else np.ascontiguousarray(qfields, dtype=np.int32)
        out = np.zeros(nq_fields * self.num_fields, dtype=np.int64)
        v = c_longlong(0)
        lib().or_pair_count_v(ctypes.byref(self._t), _dp(q), _ip(qf), q.shape[0], int(nq_fields), float(r),
                              int(nthreads), out.ctypes.data_as(POINTER(c_longlong)), ctypes.byref(v))
        return v.value, out.reshape(nq_fields, self.num_fields)

    def pair_count_brute(self, q, qfields=None, nq_fields=1, r=0.1, use_reach=True):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        qf = None if qfields is None else np.ascontiguousarray(qfields, dtype=np.int32)
        out = np.zeros(nq_fields * self.num_fields, dtype=np.int64)
        bad = lib().or_pair_count_brute(ctypes.byref(self._t), _dp(q), _ip(qf), q.shape[0],
                                        int(nq_fields), float(r), int(bool(use_reach)),
                                        out.ctypes.data_as(POINTER(c_longlong)))
        if bad:
            raise IndexError("field index outside the count matrix")
        return out.reshape(nq_fields, self.num_fields)

    def cumulative(self, q, qfields, nq_fields, radii, nthreads=8):
        """One full count per radius, as clustering._autocorr does per edge
        (corruscant/clustering.py:218-230); shape (len(radii), Nq, Nt)."""
        return np.stack([self.pair_count(q, qfields, nq_fields, r, nthreads) for r in radii])

    def __del__(self):
        try:
            lib().or_tree_free(ctypes.byref(self._t))
        except Exception:
            pass


def radec2cart64(ra, dec):
    ra = np.ascontiguousarray(ra, dtype=np.float64)
    dec = np.ascontiguousarray(dec, dtype=np.float64)
    out = np.empty((ra.size, 3), dtype=np.float64)
    lib().or_radec2cart64(_dp(ra), _dp(dec), _dp(out), ra.size)
    return out


def radecdist2cart64(ra, dec, dist):
    ra, dec, dist = (np.ascontiguousarray(a, dtype=np.float64) for a in (ra, dec, dist))
    out = np.empty((ra.size, 3), dtype=np.float64)
    lib().or_radecdist2cart64(_dp(ra), _dp(dec), _dp(dist), _dp(out), ra.size)
    return out


def radec2sph64(ra, dec):
    ra, dec = (np.ascontiguousarray(a, dtype=np.float64) for a in (ra, dec))
    th = np.empty(ra.size)
    ph = np.empty(ra.size)
    lib().or_radec2sph64(_dp(ra), _dp(dec), _dp(th), _dp(ph), ra.size)
    return th, ph


# --------------------------------------------------------------------------
# the compiled reference (oracle/_ref), when it was built here

class RefKdtree(Structure):
    """src/kdtree.h:33-41 as laid out by GCC on x86-64 (64 bytes)."""
    _fields_ = [
        ("node_data", c_void_p),
        ("size", c_int),
        ("memsize", c_int),
        ("num_fields", c_int),
        ("data", c_void_p),
        ("fields", c_void_p),
        ("args", c_void_p * 3),
    ]


assert ctypes.sizeof(RefKdtree) == 64
_ref = None


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libkdtree_ref.so"))


def ref_lib():
    global _ref
    if _ref is None:
        kd = ctypes.CDLL(os.path.join(HERE, "_ref", "libkdtree_ref.so"))
        kd.tree_construct.restype = RefKdtree
        kd.tree_construct.argtypes = [POINTER(c_double), POINTER(c_int), c_int, c_int]
        kd.pair_count.restype = POINTER(c_longlong)
        kd.pair_count.argtypes = [RefKdtree, POINTER(c_double), POINTER(c_int), c_int, c_int,
                                  c_double, c_int]
        kd.destroy.restype = None
        kd.destroy.argtypes = [RefKdtree]
        co = ctypes.CDLL(os.path.join(HERE, "_ref", "libcoords_ref.so"))
        co.radec2cart64.restype = None
        co.radec2cart64.argtypes = [POINTER(c_double)] * 3 + [c_int]
        co.radecdist2cart64.restype = None
        co.radecdist2cart64.argtypes = [POINTER(c_double)] * 4 + [c_int]
        co.radec2sph64.restype = None
        co.radec2sph64.argtypes = [POINTER(c_double)] * 4 + [c_int]
        libc = ctypes.CDLL(None)
        libc.free.argtypes = [c_void_p]
        _ref = (kd, co, libc)
    return _ref


class RefTree:
    """The reference's own tree_construct / pair_count, compiled from source."""

    def __init__(self, xyz, fields=None, num_fields=1):
        kd, _, _ = ref_lib()
        self.points = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.fields = None if fields is None else np.ascontiguousarray(fields, dtype=np.int32)
        self.num_fields = int(num_fields)
        pts = self.points
        if pts.shape[0] >= 256:
            _, cnt = np.unique(pts + 0.0, axis=0, return_counts=True)
            if cnt.max() >= 256:
                raise ValueError("reference would exit: 256 or more copies of one point")
        self.ct = kd.tree_construct(_dp(pts), _ip(self.fields), pts.shape[0], self.num_fields)
        self.size = self.ct.size
        self.memsize = self.ct.memsize
        buf = (ctypes.c_char * (40 * self.memsize)).from_address(self.ct.node_data)
        self.nodes = np.frombuffer(bytes(buf), dtype=NODE_DTYPE).copy()

    def pair_count(self, q, qfields=None, nq_fields=1, r=0.1, nthreads=1):
        kd, _, libc = ref_lib()
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        if nthreads < 1 or nthreads > q.shape[0] // 20:
            raise ValueError("reference would exit: num_threads > N/20")
        qf = None if qfields is None else np.ascontiguousarray(qfields, dtype=np.int32)
        p = kd.pair_count(self.ct, _dp(q), _ip(qf), q.shape[0], int(nq_fields), float(r), int(nthreads))
        out = np.ctypeslib.as_array(p, shape=(nq_fields * self.num_fields,)).copy()
        libc.free(ctypes.cast(p, c_void_p))
        return out.reshape(nq_fields, self.num_fields)

    def cumulative(self, q, qfields, nq_fields, radii, nthreads=1):
        return np.stack([self.pair_count(q, qfields, nq_fields, r, nthreads) for r in radii])

    def __del__(self):
        try:
            ref_lib()[0].destroy(self.ct)
        except Exception:
            pass


class RefNodes(RefTree):
    """The reference's pair_count over a given node array (memsize records in
    the src/kdtree.h layout), e.g. one built on the GPU: pair_count reads only
    node_data and num_fields (src/kdquery.c:165-175)."""

    def __init__(self, nodes, num_fields):
        ref_lib()
        self._buf = np.ascontiguousarray(nodes).view(np.uint8).copy()
        self.num_fields = int(num_fields)
        self.memsize = nodes.shape[0]
        self.ct = RefKdtree()
        self.ct.node_data = self._buf.ctypes.data
        self.ct.memsize = self.memsize
        self.ct.size = int(np.count_nonzero(nodes["weight"]))
        self.ct.num_fields = self.num_fields
        self.size = self.ct.size

    def __del__(self):
        pass  # the buffer belongs to numpy, not to the reference's destroy()


class OracleNodes(OracleTree):
    """The oracle's traversal over a given node array (no build)."""

    def __init__(self, nodes, num_fields):
        lib()
        self._buf = np.ascontiguousarray(nodes).view(np.uint8).copy()
        self.num_fields = int(num_fields)
        self.memsize = nodes.shape[0]
        self._t = OrTree()
        self._t.nodes = self._buf.ctypes.data
        self._t.memsize = self.memsize
        self._t.size = int(np.count_nonzero(nodes["weight"]))
        self._t.num_fields = self.num_fields
        self.size = self._t.size

    def __del__(self):
        pass


def ref_radec2cart64(ra, dec):
    _, co, _ = ref_lib()
    ra = np.ascontiguousarray(ra, dtype=np.float64)
    dec = np.ascontiguousarray(dec, dtype=np.float64)
    out = np.empty((ra.size, 3), dtype=np.float64)
    co.radec2cart64(_dp(ra), _dp(dec), _dp(out), ra.size)
    return out


def nodes_equal(a, b):
    """Field-by-field comparison of two node arrays (padding ignored:
    the reference copies it from malloc'd memory)."""
    if a.shape != b.shape:
        return False
    return (np.array_equal(a["value"].view(np.uint64), b["value"].view(np.uint64))
            and np.array_equal(a["field"], b["field"])
            and np.array_equal(a["weight"], b["weight"])
            and np.array_equal(a["bits"] & 0x3FF, b["bits"] & 0x3FF))
