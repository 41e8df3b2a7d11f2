"""On-disk formats around the GPU indexes.

- MMapMatrixDatabase (src/db/mmapmatrixdatabase.jl:10-283): a (dim, capacity) column-major
  matrix memory-mapped onto a file after a 4096-byte header, in the reference's own file
  format, so files written by either side open on the other.  Header (little-endian):
  magic UInt64 0x4D4D444231000001 at 0, version Int32 1 at 8, dim Int64 at 12, element
  type code Int32 at 20, logical length n Int64 at 24 (:10-31).  Capacity doubles on
  growth; n is persisted only after the data it covers is flushed (push_item!,
  append_items!: :238-283).  A Float32 file uploads to the GPU straight from the mapping.
- Index files (the role of apps/simsearch/src/persistence.jl:8-25): a SearchGraph's CSR
  adjacency (StaticAdjList's inclusive end offsets, src/adj/adjstatic.jl:27-30), hints
  and BeamSearch parameters, saved without the database; load_index reattaches one.
  The reference stores this with JLD2, which is not available here, so the file is a
  fixed binary layout of its own (save_index's docstring gives it).
- read_matrix(path, key) / split_pathkey (apps/simsearch/src/datasetio.jl:1-57) for the
  containers this image can read: .npz (key = array name), .npy (key ignored) and
  MMapMatrixDatabase files (.mmapdb, key ignored).  HDF5 and JLD2 containers need
  libraries that are not installed and raise an error that says so.
- write_results / load_results (datasetio.jl:60-110): k-NN results as .npz with keys
  "ids" (Int32, (k, m)) and optionally "dists" (Float32, (k, m)).
"""
import os
import struct

import numpy as np

from . import BeamSearch, MatrixDatabase, SearchGraph, StaticAdjList

__all__ = ["MMapMatrixDatabase", "save_index", "load_index", "read_matrix", "split_pathkey",
           "write_results", "load_results"]

MMAP_MATRIX_DB_MAGIC = 0x4D4D444231000001
MMAP_MATRIX_DB_VERSION = 1
MMAP_MATRIX_DB_HEADER_SIZE = 4096
_N_OFFSET = 24
# element type codes (mmapmatrixdatabase.jl:25-31)
_NUMTYPE_CODES = {np.dtype(np.float16): 1, np.dtype(np.float32): 2, np.dtype(np.float64): 3,
                  np.dtype(np.int8): 4, np.dtype(np.uint8): 5, np.dtype(np.int16): 6,
                  np.dtype(np.uint16): 7, np.dtype(np.int32): 8, np.dtype(np.uint32): 9,
                  np.dtype(np.int64): 10, np.dtype(np.uint64): 11}
_NUMTYPE_DECODE = {v: k for k, v in _NUMTYPE_CODES.items()}


def _fsync(f):
    f.flush()
    os.fsync(f.fileno())


class MMapMatrixDatabase(MatrixDatabase):
    """MMapMatrixDatabase(path, dim, dtype=float32, capacity_bits=8) creates a new file
    (an error if it exists); MMapMatrixDatabase(path, read_only=False) reopens one,
    taking dim, element type and length from its header.  Objects are 1-based
    (db[i]); `rows` is the (n, dim) view of the mapped data, the reference's (dim, n)."""

    def __init__(self, path, dim=None, dtype=np.float32, capacity_bits=8, read_only=False):
        self.path = str(path)
        self.closed = False
        if dim is not None:
            if os.path.exists(self.path):
                raise FileExistsError(f"MMapMatrixDatabase: file `{self.path}` already exists; "
                                      "use MMapMatrixDatabase(path) to reopen it")
            dt = np.dtype(dtype)
            if dt not in _NUMTYPE_CODES:
                raise TypeError(f"MMapMatrixDatabase: unsupported element type {dt}")
            self._dim, self._dtype, self._n, self.read_only = int(dim), dt, 0, False
            cap = 1 << int(capacity_bits)
            self._f = open(self.path, "w+b")
            self._f.truncate(MMAP_MATRIX_DB_HEADER_SIZE + self._dim * cap * dt.itemsize)
            self._f.write(struct.pack("<QiqiQ", MMAP_MATRIX_DB_MAGIC, MMAP_MATRIX_DB_VERSION,
                                      self._dim, _NUMTYPE_CODES[dt], 0))
            _fsync(self._f)
        else:
            if not os.path.isfile(self.path):
                raise FileNotFoundError(f"MMapMatrixDatabase: file `{self.path}` does not exist; "
                                        "use MMapMatrixDatabase(path, dim, dtype) to create one")
            self._f = open(self.path, "rb" if read_only else "r+b")
            head = self._f.read(32)
            if len(head) < 32:
                self._f.close()
                raise ValueError(f"MMapMatrixDatabase: `{self.path}` is not a valid file (short header)")
            magic, version, dim, code, n = struct.unpack("<QiqiQ", head)
            try:
                if magic != MMAP_MATRIX_DB_MAGIC:
                    raise ValueError(f"MMapMatrixDatabase: `{self.path}` is not a valid "
                                     "MMapMatrixDatabase file (bad magic)")
                if version != MMAP_MATRIX_DB_VERSION:
                    raise ValueError(f"MMapMatrixDatabase: unsupported format version {version}")
                if code not in _NUMTYPE_DECODE:
                    raise ValueError(f"MMapMatrixDatabase: unknown element type code {code}")
                self._dim, self._dtype, self._n = int(dim), _NUMTYPE_DECODE[code], int(n)
                cap = self._capacity_on_disk()
                if cap < self._n:
                    raise ValueError(f"MMapMatrixDatabase: corrupt file `{self.path}`: physical "
                                     f"capacity ({cap}) is smaller than the recorded length ({n})")
            except Exception:
                self._f.close()
                raise
            self.read_only = bool(read_only)
        self._map()

    # ------------------------------------------------------------------ mapping
    def _capacity_on_disk(self):
        size = os.path.getsize(self.path) - MMAP_MATRIX_DB_HEADER_SIZE
        return max(0, size // (self._dim * self._dtype.itemsize))

    def _map(self):
        cap = self._capacity_on_disk()
        self._data = np.memmap(self._f, dtype=self._dtype, mode="r" if self.read_only else "r+",
                               offset=MMAP_MATRIX_DB_HEADER_SIZE, shape=(max(cap, 1), self._dim))
        self._cap = cap

    def _grow(self, mincap):
        newcap = max(1, self._cap)
        while newcap < mincap:
            newcap *= 2
        self._data.flush()
        self._f.truncate(MMAP_MATRIX_DB_HEADER_SIZE + self._dim * newcap * self._dtype.itemsize)
        self._map()

    def _persist_n(self, n):
        self._f.seek(_N_OFFSET)
        self._f.write(struct.pack("<q", n))
        _fsync(self._f)

    def _writable(self, what):
        if self.closed:
            raise ValueError("MMapMatrixDatabase: the database is closed")
        if self.read_only:
            raise PermissionError(f"MMapMatrixDatabase: cannot {what} on a read_only database")

    # ------------------------------------------------------------------ database API
    @property
    def rows(self):
        return self._data[:self._n]

    @property
    def data(self):
        """The whole mapped capacity, (capacity, dim) rows."""
        return self._data

    @property
    def dtype(self):
        return self._dtype

    @property
    def capacity(self):
        return self._cap

    def __setitem__(self, i, value):
        self._writable("setindex!")
        if not 1 <= i <= self._n:
            raise IndexError(i)
        self._data[i - 1] = value
        self.touch()

    def push_item(self, v):
        """push_item!: write v, flush it, then persist the new length (:247-256)."""
        self._writable("push_item!")
        n = self._n + 1
        if n > self._cap:
            self._grow(n)
        self._data[n - 1] = np.asarray(v, self._dtype)
        self._data.flush()
        self._n = n
        self._persist_n(n)
        self.touch()
        return self

    def append_items(self, B):
        """append_items!: write every object of B (an iterable of vectors, or an (m, dim)
        array of rows), flush once, then persist the new length once (:270-283)."""
        self._writable("append_items!")
        n0 = n = self._n
        if isinstance(B, np.ndarray) and B.ndim == 2:
            m = B.shape[0]
            if m and n + m > self._cap:
                self._grow(n + m)
            self._data[n:n + m] = B
            n += m
        else:
            for v in B:
                n += 1
                if n > self._cap:
                    self._grow(n)
                self._data[n - 1] = np.asarray(v, self._dtype)
        if n > n0:
            self._data.flush()
            self._n = n
            self._persist_n(n)
            self.touch()
        return self

    def close(self):
        if not self.closed:
            self.closed = True
            if not self.read_only:
                self._data.flush()
            del self._data
            self._f.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __repr__(self):
        return (f"MMapMatrixDatabase{{{self._dim},{self._dtype}}}(path={self.path!r}, "
                f"length={self._n}, capacity={self._cap}, read_only={self.read_only})")


# ---------------------------------------------------------------------- index files
_INDEX_MAGIC = b"SSGIDX01"
_INDEX_HEAD = struct.Struct("<8sqqqqifqi")   # magic, n, nnz, nhints, dim, bsize, delta, maxvisits, metric


def save_index(path, G):
    """save_index(path, G): write SearchGraph G without its database.  Layout
    (little-endian): 8-byte magic "SSGIDX01", Int64 n, nnz, nhints, dim, Int32 bsize,
    Float32 delta, Int64 maxvisits, Int32 metric code; then Int64 offsets[n] (inclusive
    end offsets), UInt32 end_point[nnz] (1-based) and UInt32 hints[nhints]."""
    from . import _metric
    if not isinstance(G, SearchGraph):
        raise TypeError("save_index expects a SearchGraph")
    off = np.ascontiguousarray(G.adj.offset, np.int64)
    ep = np.ascontiguousarray(G.adj.end_point, np.uint32)
    hints = np.ascontiguousarray(G.hints, np.uint32)
    a = G.algo
    tmp = str(path) + ".tmp"
    with open(tmp, "wb") as f:
        f.write(_INDEX_HEAD.pack(_INDEX_MAGIC, len(off), len(ep), len(hints), G.db.dim, int(a.bsize),
                                 float(a.delta), int(a.maxvisits), _metric(G.dist)))
        f.write(off.tobytes())
        f.write(ep.tobytes())
        f.write(hints.tobytes())
        _fsync(f)
    os.replace(tmp, path)


def load_index(path, dist, db, device=0):
    """load_index(path, dist, db): the SearchGraph saved at path over database db (which
    must hold as many objects, of the same dimension, as the saved graph's)."""
    from . import _as_db, _metric
    db = _as_db(db)
    with open(path, "rb") as f:
        head = f.read(_INDEX_HEAD.size)
        if len(head) < _INDEX_HEAD.size:
            raise ValueError(f"load_index: `{path}` is truncated")
        magic, n, nnz, nh, dim, bsize, delta, maxvisits, metric = _INDEX_HEAD.unpack(head)
        if magic != _INDEX_MAGIC:
            raise ValueError(f"load_index: `{path}` is not an index file (bad magic)")
        if n != len(db) or dim != db.dim:
            raise ValueError(f"load_index: the index holds {n} objects of dimension {dim}; the "
                             f"database has {len(db)} of dimension {db.dim}")
        if metric != _metric(dist):
            raise ValueError("load_index: the distance differs from the one the index was saved with")
        off = np.fromfile(f, np.int64, n)
        ep = np.fromfile(f, np.uint32, nnz)
        hints = np.fromfile(f, np.uint32, nh)
    if len(off) != n or len(ep) != nnz or len(hints) != nh:
        raise ValueError(f"load_index: `{path}` is truncated")
    if n and (off[-1] != nnz or (np.diff(off) < 0).any() or off[0] < 0):
        raise ValueError(f"load_index: `{path}` has inconsistent offsets")
    if nnz and (ep.min() < 1 or ep.max() > n):
        raise ValueError(f"load_index: `{path}` has neighbour ids out of range")
    return SearchGraph(dist, db, StaticAdjList(off, ep), hints, BeamSearch(bsize, delta, maxvisits),
                       device=device)


# ---------------------------------------------------------------------- datasets and results
def split_pathkey(spec, require_key=True):
    """split_pathkey(spec; require_key) -> (path, key) (datasetio.jl:1-17): split on the
    last ':'; without one, an error when a key is required, else key None; an empty key
    is an error."""
    spec = str(spec)
    i = spec.rfind(":")
    if i < 0:
        if require_key:
            raise ValueError(f"expected 'path:key' form, got '{spec}' (no ':' found)")
        return spec, None
    path, key = spec[:i], spec[i + 1:]
    if not key:
        raise ValueError(f"empty key after ':' in '{spec}'")
    return path, key


def read_matrix(path, key=None):
    """The (dim, n) Float32 matrix stored under key in path, as a MatrixDatabase-ready
    array (datasetio.jl:57).  .npz: the array named key, read as (dim, n) columns like
    the reference's containers; .npy: the file's array; .mmapdb: an MMapMatrixDatabase
    file's rows (returned as (dim, n))."""
    ext = os.path.splitext(str(path))[1].lower()
    if ext == ".npz":
        with np.load(path, allow_pickle=False) as z:
            if key not in z.files:
                raise KeyError(f"key '{key}' not found in '{path}' (has {z.files})")
            return np.asarray(z[key], np.float32)
    if ext == ".npy":
        return np.asarray(np.load(path, allow_pickle=False), np.float32)
    if ext == ".mmapdb":
        db = MMapMatrixDatabase(path, read_only=True)
        try:
            return np.asarray(db.rows, np.float32).T.copy()
        finally:
            db.close()
    if ext in (".h5", ".hdf5", ".jld2"):
        raise RuntimeError(f"'{path}': HDF5/JLD2 containers need h5py, which is not installed; "
                           "convert the matrix to .npz or an MMapMatrixDatabase file")
    raise ValueError(f"unsupported extension '{ext}' for '{path}' (expected .npz, .npy or .mmapdb)")


def write_results(path, ids, dists=None):
    """write_results(path, ids, dists) -> a .npz file with keys "ids" ((k, m) Int32) and,
    if given, "dists" ((k, m) Float32) (datasetio.jl:60-85)."""
    if not str(path).lower().endswith(".npz"):
        raise ValueError(f"results path must end in .npz, got '{path}'")
    out = {"ids": np.asarray(ids, np.int32)}
    if dists is not None:
        out["dists"] = np.asarray(dists, np.float32)
    with open(path, "wb") as f:
        np.savez(f, **out)


def load_results(spec):
    """load_results(spec) -> (ids Int32 (k, m), dists Float32 (k, m) or None): a file
    written by write_results, or "path:key" naming an ids array with an optional sibling
    "<key>_dists" (datasetio.jl:87-110)."""
    path, key = split_pathkey(spec, require_key=False)
    with np.load(path, allow_pickle=False) as z:
        ik, dk = ("ids", "dists") if key is None else (key, key + "_dists")
        if ik not in z.files:
            raise KeyError(f"key '{ik}' not found in '{path}'")
        ids = np.asarray(z[ik], np.int32)
        dists = np.asarray(z[dk], np.float32) if dk in z.files else None
    return ids, dists
