"""Distributed algebra objects returned by the assembly (mirror of the
PartitionedArrays types the reference returns, src/Algebra.jl:1013-1015).

* ``SubAssembledRows`` / ``FullyAssembledRows`` -- src/Algebra.jl:397-398
* ``LocalIndices`` / ``PRange``   -- OwnAndGhostIndices built at src/Algebra.jl:1240-1252
* ``PSparseMatrix``              -- fields matrix_partition, row_partition, col_partition
                                    (test/FESpacesTests.jl:29-31)
* ``PVector``                    -- vector on the row partition of A (src/Algebra.jl:720-753),
                                    with PArrays' consistent! and assemble!
* ``DevicePVector``              -- the same layout in device memory, with consistent!,
                                    assemble!, axpy! (src/Algebra.jl:42-49), change_ghost
                                    (src/Algebra.jl:270-290) and mul!(y, A, x) on the device

Local matrices are 0-based CSR (the SparseMatrixCSR{0,Float64,Int} layout of
test/PoissonTests.jl:40); ``to_csc()`` gives the default SparseMatrixCSC view.
"""
import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib


class SubAssembledRows:
    """Integrate owned cells, send ghost-row contributions to their owners."""

    code = 0


class FullyAssembledRows:
    """Integrate owned and ghost cells, drop ghost rows: no matrix exchange."""

    code = 1


class LocalIndices:
    """Own-then-ghost local indices of one part (0-based global ids)."""

    def __init__(self, part, ngids, own_to_global, ghost_to_global, ghost_to_owner):
        self.part = part
        self.ngids = int(ngids)
        self.own_to_global = np.asarray(own_to_global, np.int64)
        self.ghost_to_global = np.asarray(ghost_to_global, np.int64)
        self.ghost_to_owner = np.asarray(ghost_to_owner, np.int32)

    def own_length(self):
        return len(self.own_to_global)

    def ghost_length(self):
        return len(self.ghost_to_global)

    def local_length(self):
        return self.own_length() + self.ghost_length()

    def local_to_global(self):
        return np.concatenate([self.own_to_global, self.ghost_to_global])

    def local_to_owner(self):
        return np.concatenate([np.full(self.own_length(), self.part, np.int32), self.ghost_to_owner])

    def global_to_local(self):
        m = {}
        for lid, g in enumerate(self.local_to_global()):
            m[int(g)] = lid
        return m


class PRange:
    def __init__(self, partition):
        self.partition = list(partition)

    def __len__(self):
        return self.partition[0].ngids if self.partition else 0


class PVector:
    """Per-part local vectors (own then ghost entries) on an index partition."""

    def __init__(self, values, index_partition):
        self.partition = [np.asarray(v) for v in values]
        self.index_partition = list(index_partition)

    def own_values(self):
        return [v[: ip.own_length()] for v, ip in zip(self.partition, self.index_partition)]

    def to_global(self):
        n = self.index_partition[0].ngids
        out = np.zeros(n)
        for v, ip in zip(self.partition, self.index_partition):
            out[ip.own_to_global] = v[: ip.own_length()]
        return out

    def __len__(self):
        return self.index_partition[0].ngids

    def copy(self):
        return PVector([v.copy() for v in self.partition], self.index_partition)

    def _own_global(self):
        g = np.zeros(self.index_partition[0].ngids)
        for v, ip in zip(self.partition, self.index_partition):
            g[ip.own_to_global] = v[: ip.own_length()]
        return g

    def consistent(self):
        """consistent!: every ghost entry takes its owner's value."""
        g = self._own_global()
        for v, ip in zip(self.partition, self.index_partition):
            v[ip.own_length():] = g[ip.ghost_to_global]
        return self

    def assemble(self):
        """assemble!: ghost entries are added into their owner's entry (the owner's
        value first, then the copies in ascending part and local index), then zeroed."""
        g = self._own_global()
        for v, ip in zip(self.partition, self.index_partition):
            np.add.at(g, ip.ghost_to_global, v[ip.own_length():])  # sequential, in index order
        for v, ip in zip(self.partition, self.index_partition):
            v[: ip.own_length()] = g[ip.own_to_global]
            v[ip.own_length():] = 0.0
        return self

    def concatenated(self):
        """Every part's local vector, concatenated in part order (the DevicePVector layout)."""
        return np.concatenate(self.partition) if self.partition else np.zeros(0)


class DevicePVector:
    """A PVector in device memory on a problem's row (default) or column PRange:
    every part's local vector (own entries, then ghost entries) concatenated in
    part order, owned by the library and tied to its problem."""

    def __init__(self, problem, which=_lib.PRANGE_ROWS):
        h = ctypes.c_void_p()
        _lib.call("gdx_pvector_create", problem.handle, int(which), ctypes.byref(h))
        self.handle, self.problem, self.which = h, problem, int(which)
        n = ctypes.c_int64()
        _lib.call("gdx_pvector_length", h, ctypes.byref(n))
        self.n = n.value

    def set(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != (self.n,):
            raise ValueError(f"a PVector on this PRange has {self.n} local entries")
        _lib.call("gdx_pvector_set", self.handle, v.ctypes.data)
        return self

    def get(self):
        out = np.empty(self.n)
        _lib.call("gdx_pvector_get", self.handle, out.ctypes.data)
        return out

    def fill(self, value):
        _lib.call("gdx_pvector_fill", self.handle, float(value))
        return self

    def copy_rhs(self):
        """The assembled b (row PRange)."""
        _lib.call("gdx_pvector_copy_rhs", self.handle)
        return self

    def consistent(self):
        _lib.call("gdx_pvector_consistent", self.handle)
        return self

    def assemble(self):
        _lib.call("gdx_pvector_assemble", self.handle)
        return self

    def axpy(self, alpha, x):
        """axpy!(alpha, x, self): self = alpha x + self, then consistent!."""
        _lib.call("gdx_pvector_axpy", float(alpha), x.handle, self.handle)
        return self

    def change_ghost_from(self, src, make_consistent=False):
        """change_ghost: src's own values on this PVector's PRange, ghosts zero."""
        _lib.call("gdx_pvector_change_ghost", src.handle, self.handle, int(bool(make_consistent)))
        return self

    def close(self):
        if self.handle:
            _lib.call("gdx_pvector_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PSparseMatrix:
    """Per-part local CSR matrices on row/col index partitions."""

    def __init__(self, matrix_partition, row_partition, col_partition):
        self.matrix_partition = list(matrix_partition)
        self.row_partition = list(row_partition)
        self.col_partition = list(col_partition)

    @property
    def shape(self):
        n = self.row_partition[0].ngids
        m = self.col_partition[0].ngids
        return (n, m)

    def to_global(self):
        """Owned rows of every part as one global scipy CSR matrix."""
        I, J, V = [], [], []
        for A, rows, cols in zip(self.matrix_partition, self.row_partition, self.col_partition):
            no = rows.own_length()
            Ao = A[:no].tocoo()
            I.append(rows.own_to_global[Ao.row])
            J.append(cols.local_to_global()[Ao.col])
            V.append(Ao.data)
        return sp.csr_matrix((np.concatenate(V), (np.concatenate(I), np.concatenate(J))), shape=self.shape)

    def to_csc(self):
        return [A.tocsc() for A in self.matrix_partition]

    def copy(self):
        """Base.copy(::PSparseMatrix) (src/Algebra.jl:34-38): values copied, partitions shared."""
        return PSparseMatrix([A.copy() for A in self.matrix_partition], self.row_partition, self.col_partition)

    def __matmul__(self, x):
        """A*x for a global numpy vector or a PVector (owned rows)."""
        xg = x.to_global() if isinstance(x, PVector) else np.asarray(x)
        return self.to_global() @ xg
