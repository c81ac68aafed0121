"""Multi-field spaces and block assembly (mirror of src/MultiField.jl:470-560 and
src/BlockPartitionedArrays.jl).

``MultiFieldFESpace(spaces, style=BlockMultiFieldStyle())`` over scalar
Lagrangian spaces of one model; a block weak form is a list of terms
``Block(i, j, Diffusion(kappa) | Mass(sigma), dOmega)`` (test field i, trial
field j: int(kappa grad(u_j).grad(v_i)) or int(sigma u_j v_i)) and one source
per test field, as in

  biform((u1,u2),(v1,v2)) = int( grad(u1).grad(v1) + u2 v2 + u1 v2 ) dOmega
  liform((v1,v2))          = int( v1 + v2 ) dOmega

of test/BlockSparseMatrixAssemblersTests.jl:110-114.  Block (i, j) is
assembled on the GPU as its own problem on the pair of spaces (rows: field i's
dof PRange, columns: field j's), diffusion and reaction weights on the
mapped-grid path (gdx_set_reaction); the block vector of field i is its
source integral minus the liftings of every block in its row, the RHS the
reference's block assembly writes.  The result is a ``BlockPMatrix`` (blocks
without terms are None, as Gridap leaves them empty) and a ``BlockPVector``.
"""
from .algebra import FullyAssembledRows, PVector
from .fespaces import (ConstantCoefficient, ConstantSource, Diffusion, FourierExpCoefficient, Integral,
                       SparseMatrixAssembler, Source, TrialFESpace, collect_cell_matrix_and_vector)


class BlockMultiFieldStyle:
    """Gridap's BlockMultiFieldStyle(): one block per field."""


class MultiFieldFESpace:
    """MultiFieldFESpace(spaces; style) of scalar spaces on one model."""

    def __init__(self, spaces, style=None):
        self.spaces = list(spaces)
        self.style = style if style is not None else BlockMultiFieldStyle()
        if not isinstance(self.style, BlockMultiFieldStyle):
            raise NotImplementedError("only BlockMultiFieldStyle() is lowered")
        models = {id(s.model) for s in self.spaces}
        if len(models) != 1:
            raise ValueError("the fields must live on one model")

    def num_fields(self):
        return len(self.spaces)


class Mass:
    """sigma u v (a constant sigma)."""

    def __init__(self, sigma=1.0):
        self.sigma = float(sigma)


class Block:
    """A term of the block weak form: test field i, trial field j."""

    def __init__(self, i, j, integrand, measure):
        if not isinstance(integrand, (Diffusion, Mass)):
            raise NotImplementedError("block terms are Diffusion or Mass")
        self.i, self.j, self.integrand, self.measure = int(i), int(j), integrand, measure


class BlockPMatrix:
    """blocks[i][j]: PSparseMatrix or None."""

    def __init__(self, blocks):
        self.blocks = blocks

    def nblocks(self):
        return len(self.blocks), len(self.blocks[0]) if self.blocks else 0


class BlockPVector:
    def __init__(self, blocks):
        self.blocks = blocks


def _space(field):
    return field.space if isinstance(field, TrialFESpace) else field


def assemble_block_matrix_and_vector(X, Y, terms, sources, strategy=None):
    """Block assembly of the weak form ``terms`` (list of Block) with the linear
    form ``sources`` (one Source per test field, None for none) on the trial
    space X and the test space Y.  Returns (BlockPMatrix, BlockPVector, problems);
    problems[(i, j)] is (assembler, GpuProblem) of block (i, j) (its device
    matrix and RHS stay usable for solves)."""
    strategy = strategy if strategy is not None else FullyAssembledRows()
    nf = Y.num_fields()
    if X.num_fields() != nf or len(sources) != nf:
        raise ValueError("X, Y and the sources must have one entry per field")
    # the weights of each block: (kappa, diffusion weight, reaction)
    spec = {}
    for t in terms:
        if not (0 <= t.i < nf and 0 <= t.j < nf):
            raise ValueError("block index out of range")
        kap, dw, rw, meas = spec.get((t.i, t.j), (None, 0.0, 0.0, t.measure))
        if t.measure.degree != meas.degree:
            raise NotImplementedError("the terms of a block must use one Measure")
        if isinstance(t.integrand, Diffusion):
            if kap is not None:
                raise NotImplementedError("one diffusion term per block")
            kap, dw = t.integrand.kappa, 1.0
        else:
            rw += t.integrand.sigma
        spec[(t.i, t.j)] = (kap, dw, rw, meas)
    rows = [sorted(j for (ii, j) in spec if ii == i) for i in range(nf)]
    # everything is checked before any block is assembled
    for i, row in enumerate(rows):
        if not row:
            raise ValueError(f"test field {i} has no term")
        for j in row:
            kap, dw, rw, meas = spec[(i, j)]
            if isinstance(kap, FourierExpCoefficient) and rw:
                raise NotImplementedError("a reaction term beside a Fourier kappa is not lowered")
            U, V = _space(X.spaces[j]), Y.spaces[i]
            if U is not V and (U.reffe.order != V.reffe.order or U.dirichlet != V.dirichlet):
                raise NotImplementedError("the fields of a block share one Lagrangian space")
    blocks = [[None] * nf for _ in range(nf)]
    rhs = [None] * nf
    probs = {}
    for i in range(nf):
        row = rows[i]
        for j in row:
            kap, dw, rw, meas = spec[(i, j)]
            kappa = kap if kap is not None else ConstantCoefficient(1.0)
            # the field's source goes with the first block of its row
            src = sources[i] if (j == row[0] and sources[i] is not None) else Source(ConstantSource(0.0))
            U, V = X.spaces[j], Y.spaces[i]
            assem = SparseMatrixAssembler(U, V, strategy)
            data = collect_cell_matrix_and_vector(U, V, Integral(Diffusion(kappa), meas), Integral(src, meas), None)
            prob = assem._problem_for(data)
            if (dw, rw) != (1.0, 0.0):
                if not prob.mapped_coords_set:
                    prob.set_vertex_coords(V.model.uniform_vertices())
                prob.set_reaction(dw, rw)
            prob.numeric()
            prob.sync()
            A, b = assem.export()
            blocks[i][j] = A
            # every block of row i has field i's rows: the vectors add part by part
            rhs[i] = b if rhs[i] is None else PVector([x + y for x, y in zip(rhs[i].partition, b.partition)],
                                                      rhs[i].index_partition)
            probs[(i, j)] = (assem, prob)
    return BlockPMatrix(blocks), BlockPVector(rhs), probs
