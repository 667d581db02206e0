"""The reference's matrix text format (src/read_write.zig:24-80), for exchanging fixtures with a Zinq
binary: a header line "nrow ncol", then one line per row with every element printed as {d:20.14}
(fixed point, 14 fractional digits, right-aligned in 20 columns) and separated by one space.  Tensors
are written through asMatrix (src/tensor.zig:258-271): the first ceil(N/2) indices fold into the row,
so the dense ERIs g[p,q,r,s] travel as an n^2 x n^2 matrix (src/molecular_integrals.zig:322-327).
"""
import numpy as np


def write_matrix(path, A):
    A = np.asarray(A, dtype=np.float64)
    if A.ndim != 2:
        raise ValueError("write_matrix takes a 2-D array; use write_tensor for more indices")
    nrow, ncol = A.shape
    with open(path, "w") as f:
        f.write(f"{nrow} {ncol}\n")
        for row in A:
            f.write(" ".join(f"{x:20.14f}" for x in row))
            f.write("\n")


def read_matrix(path):
    """Header "nrow ncol", then nrow*ncol whitespace-separated numbers (line breaks not significant,
    as readMatrix fills the data in order)."""
    with open(path) as f:
        head = f.readline().split()
        if len(head) < 2:
            raise ValueError(f"{path}: bad header")
        nrow, ncol = int(head[0]), int(head[1])
        data = np.array(f.read().split(), dtype=np.float64)
    if data.size != nrow * ncol:
        raise ValueError(f"{path}: {data.size} elements for a {nrow} x {ncol} matrix")
    return data.reshape(nrow, ncol)


def as_matrix(T):
    """asMatrix of src/tensor.zig:258-271: rows = product of the first ceil(N/2) extents."""
    T = np.asarray(T)
    k = (T.ndim + 1) // 2
    return T.reshape(int(np.prod(T.shape[:k])), -1)


def write_tensor(path, T):
    write_matrix(path, as_matrix(T))


def read_coulomb(path):
    """Dense ERIs g[p,q,r,s] = <pq|rs> from an n^2 x n^2 coulomb file."""
    A = read_matrix(path)
    n = int(round(A.shape[0] ** 0.5))
    if n * n != A.shape[0] or A.shape[0] != A.shape[1]:
        raise ValueError(f"{path}: not an n^2 x n^2 matrix")
    return A.reshape(n, n, n, n)
