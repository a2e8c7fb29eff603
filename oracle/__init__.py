'''
The oracle: a CPU restatement of the reference engine's block-sparse contraction path
(categories, legs, fusion trees, layouts, leg moves, contraction, einsum, splittings, Lanczos, MPS,
MPO, Heisenberg and the two-site DMRG driver), in Python and numpy.

TEST INFRASTRUCTURE ONLY.  Only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of
`bench.py` may import it, and only as the checker; the product (`tentorch_amd/`) never imports,
links or runs anything from here, and fails loudly when its HIP extension is missing.

Pinning (what each part is checked against, beyond its own restatement):
- exact diagonalisation of the Heisenberg chain (known answers E0(L=4) = -(3+2*sqrt(3))/4, L=6, 8,
  10, 12) through the whole DMRG driver;
- for abelian categories, a dense restatement: a block-sparse tensor expands to a dense array through
  the row order of its fusion trees, and `place`/`contract`/`einsum` must equal
  transpose/tensordot/numpy.einsum on that array (oracle/dense.py);
- SU(2) and Fibonacci F/R: unitarity, pentagon and hexagon identities, and the closed forms;
- place invariants for non-abelian categories: norm preserved, place followed by its inverse is the
  identity, two chains to the same layout agree.
Parity unpinned: the bend coefficient gauge of non-abelian categories beyond these invariants, and
the tie order of `std::sort` in the SVD ranking (`svd.cpp:79-80`).
'''
