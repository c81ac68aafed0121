"""Distributed algebra objects returned by the assembly (mirror of the
PartitionedArrays types the reference returns, src/Algebra.jl:1013-1015).

* ``SubAssembledRows`` / ``FullyAssembledRows`` -- src/Algebra.jl:397-398
* ``LocalIndices`` / ``PRange``   -- OwnAndGhostIndices built at src/Algebra.jl:1240-1252
* ``PSparseMatrix``              -- fields matrix_partition, row_partition, col_partition
                                    (test/FESpacesTests.jl:29-31)
* ``PVector``                    -- vector on the row partition of A (src/Algebra.jl:720-753)

Local matrices are 0-based CSR (the SparseMatrixCSR{0,Float64,Int} layout of
test/PoissonTests.jl:40); ``to_csc()`` gives the default SparseMatrixCSC view.
"""
import numpy as np
import scipy.sparse as sp


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
