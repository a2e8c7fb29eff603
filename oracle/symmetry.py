'''
Oracle: the fusion categories, legs, fusion trees and block layouts of the reference engine,
restated on the CPU in plain Python and numpy.

TEST INFRASTRUCTURE ONLY.  Nothing under `oracle/` is part of the product: only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import it, and then only as
the checker.  The product path lives in `tentorch_amd/` and never imports this package.

Each function cites the reference file and line it restates (paths relative to the reference
checkout).  Sectors are tuples of ints, ordered as Python orders tuples, which is the order of
`SectorOf::operator<` (`tentorch/symmetry/cat/cat.h:118-128`).
'''

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

Sector = Tuple[int, ...]


def render(sector: Sector) -> str:
    '''`render` (`tentorch/symmetry/tree/tree.cpp:16-31`).'''
    if len(sector) == 1:
        return f'({sector[0]},)'
    return '(' + ', '.join(str(v) for v in sector) + ')'


class Category:
    '''
    `FusionCategory` (`tentorch/symmetry/cat/cat.h:216-429`, `cat.cpp:29-109`): the outcome table
    with its cache, `n_symbol`, and the F/R addressing `f_rows`/`f_columns`/`r_rows`.
    '''

    name = 'Category'
    abelian = False

    def __init__(self) -> None:
        self._outcomes: Dict[Tuple[Sector, Sector], List[Tuple[Sector, int]]] = {}
        self._f: Dict[tuple, np.ndarray] = {}
        self._r: Dict[tuple, np.ndarray] = {}

    # -- interface a concrete category implements
    def compute_outcomes(self, a: Sector, b: Sector) -> List[Tuple[Sector, int]]:
        raise NotImplementedError

    def compute_F(self, a, b, c, d) -> np.ndarray:
        raise NotImplementedError

    def compute_R(self, b, c) -> np.ndarray:
        raise NotImplementedError

    def q_dim(self, a: Sector) -> float:
        raise NotImplementedError

    def dual(self, a: Sector) -> Sector:
        raise NotImplementedError

    def trivial(self) -> Sector:
        return (0,)

    def frobenius_schur_sign(self, a: Sector) -> float:
        return 1.0

    # -- shared (cat.cpp:29-109)
    def outcomes(self, a: Sector, b: Sector) -> List[Tuple[Sector, int]]:
        key = (a, b)
        found = self._outcomes.get(key)
        if found is None:
            found = self.compute_outcomes(a, b)
            self._outcomes[key] = found
        return found

    def n_symbol(self, a: Sector, b: Sector, c: Sector) -> int:
        for sector, mult in self.outcomes(a, b):
            if sector == c:
                return mult
        return 0

    def can_fuse_to(self, a, b, c) -> bool:
        return self.n_symbol(a, b, c) > 0

    def f_rows(self, a, b, c, d):
        '''`f_rows` (`cat.cpp:62-78`): (e, inner, outer) for e in outcomes(a, b).'''
        rows = []
        for e, mult in self.outcomes(a, b):
            copies = self.n_symbol(e, c, d)
            for inner in range(mult):
                for outer in range(copies):
                    rows.append((e, inner, outer))
        return rows

    def f_columns(self, a, b, c, d):
        '''`f_columns` (`cat.cpp:80-96`): (f, inner, outer) for f in outcomes(b, c).'''
        cols = []
        for f, mult in self.outcomes(b, c):
            copies = self.n_symbol(a, f, d)
            for inner in range(mult):
                for outer in range(copies):
                    cols.append((f, inner, outer))
        return cols

    def r_rows(self, b, c):
        '''`r_rows` (`cat.cpp:98-109`).'''
        rows = []
        for e, mult in self.outcomes(b, c):
            for copy in range(mult):
                rows.append((e, copy))
        return rows

    def F_symbol(self, a, b, c, d) -> np.ndarray:
        key = (a, b, c, d)
        found = self._f.get(key)
        if found is None:
            found = np.asarray(self.compute_F(a, b, c, d), dtype=np.complex128)
            self._f[key] = found
        return found

    def R_symbol(self, b, c) -> np.ndarray:
        key = (b, c)
        found = self._r.get(key)
        if found is None:
            found = np.asarray(self.compute_R(b, c), dtype=np.complex128)
            self._r[key] = found
        return found

    def clear_cache(self) -> None:
        self._outcomes.clear()
        self._f.clear()
        self._r.clear()

    def __repr__(self) -> str:
        return self.name


class Group(Category):
    '''
    `Group` (`tentorch/symmetry/cat/group.cpp:33-130`): F and R as overlaps of Clebsch-Gordan
    tensors divided by the quantum dimension of the coupled sector.
    '''

    def __init__(self) -> None:
        super().__init__()
        self._cg: Dict[tuple, np.ndarray] = {}

    def compute_cg(self, a, b, c, copy) -> np.ndarray:
        raise NotImplementedError

    def clebsch_gordan(self, a, b, c, copy) -> np.ndarray:
        key = (a, b, c, copy)
        found = self._cg.get(key)
        if found is None:
            found = np.asarray(self.compute_cg(a, b, c, copy), dtype=np.complex128)
            self._cg[key] = found
        return found

    def compute_F(self, a, b, c, d):
        '''`Group::compute_F_symbol` (`group.cpp:44-83`).'''
        left = [np.einsum('abe,ecd->abcd', self.clebsch_gordan(a, b, e, inner),
                          self.clebsch_gordan(e, c, d, outer))
                for e, inner, outer in self.f_rows(a, b, c, d)]
        right = [np.einsum('bce,aed->abcd', self.clebsch_gordan(b, c, f, inner),
                           self.clebsch_gordan(a, f, d, outer))
                 for f, inner, outer in self.f_columns(a, b, c, d)]
        if not left or not right:
            return np.zeros((len(left), len(right)), dtype=np.complex128)
        lm = np.stack(left).reshape(len(left), -1)
        rm = np.stack(right).reshape(len(right), -1)
        return (lm @ rm.conj().T) / self.q_dim(d)

    def compute_R(self, b, c):
        '''`Group::compute_R_symbol` (`group.cpp:93-130`).'''
        rows = self.r_rows(b, c)
        size = len(rows)
        value = np.zeros((size, size), dtype=np.complex128)
        start = 0
        while start < size:
            e = rows[start][0]
            end = start
            while end < size and rows[end][0] == e:
                end += 1
            copies = end - start
            if self.n_symbol(c, b, e) != copies:
                raise RuntimeError('R_symbol needs N^e_bc == N^e_cb for every outcome e of b (x) c')
            straight = np.stack([self.clebsch_gordan(b, c, e, k) for k in range(copies)])
            swapped = np.stack([self.clebsch_gordan(c, b, e, k).transpose(1, 0, 2)
                                for k in range(copies)])
            block = (straight.reshape(copies, -1) @ swapped.reshape(copies, -1).conj().T) / self.q_dim(e)
            value[start:end, start:end] = block
            start = end
        return value


class AbelianGroup(Group):
    '''`AbelianGroup` (`tentorch/symmetry/cat/abelian/abelian.cpp:34-58`).'''

    abelian = True

    def combine(self, a, b) -> Sector:
        raise NotImplementedError

    def inverse(self, a) -> Sector:
        raise NotImplementedError

    def compute_outcomes(self, a, b):
        return [(self.combine(a, b), 1)]

    def dual(self, a):
        return self.inverse(a)

    def q_dim(self, a):
        return 1.0

    def compute_cg(self, a, b, c, copy):
        if copy != 0 or self.combine(a, b) != c:
            raise RuntimeError('an abelian fusion couples once, to the combined sector')
        return np.ones((1, 1, 1))

    # An abelian F/R is 1 whenever it exists; computed in closed form rather than by the overlap so
    # the oracle stays fast.  The overlap form above gives the same value (checked in the tests).
    def compute_F(self, a, b, c, d):
        side = 1 if self.combine(self.combine(a, b), c) == d else 0
        return np.ones((side, side), dtype=np.complex128)

    def compute_R(self, b, c):
        return np.ones((1, 1), dtype=np.complex128)


class NoSymmetry(AbelianGroup):
    '''`NoSymmetry` (`abelian.cpp:60-64`).'''
    name = 'NoSymmetry'

    def combine(self, a, b):
        return (0,)

    def inverse(self, a):
        return (0,)


class U1Symmetry(AbelianGroup):
    '''`U1Symmetry` (`abelian.cpp:66-72`): combine adds, inverse negates.'''
    name = 'U(1)'

    def combine(self, a, b):
        return (a[0] + b[0],)

    def inverse(self, a):
        return (-a[0],)


class ZNSymmetry(AbelianGroup):
    '''`ZNSymmetry` (`abelian.cpp:74-81`): residues in [0, N).'''

    def __init__(self, modulus: int) -> None:
        if modulus <= 0:
            raise ValueError(f'The modulus of Z_N must be positive, not {modulus}.')
        super().__init__()
        self.modulus = modulus
        self.name = f'Z_{modulus}'

    def combine(self, a, b):
        return ((a[0] + b[0]) % self.modulus,)

    def inverse(self, a):
        return ((-a[0]) % self.modulus,)


class FermionParity(Category):
    '''`FermionParity` (`abelian.cpp:83-115`): Koszul sign -1 for odd x odd.'''
    name = 'FermionParity'
    abelian = True

    def compute_outcomes(self, a, b):
        return [(((a[0] + b[0]) % 2,), 1)]

    def compute_F(self, a, b, c, d):
        side = 1 if (a[0] + b[0] + c[0]) % 2 == d[0] else 0
        return np.ones((side, side), dtype=np.complex128)

    def compute_R(self, b, c):
        return np.full((1, 1), -1.0 if (b[0] == 1 and c[0] == 1) else 1.0, dtype=np.complex128)

    def q_dim(self, a):
        return 1.0

    def dual(self, a):
        return a

    def frobenius_schur_sign(self, a):
        return -1.0 if a[0] == 1 else 1.0


def su2_coefficients(j1: int, j2: int) -> Dict[int, np.ndarray]:
    '''
    Clebsch-Gordan tensors of SU(2) by lowering highest weights, labels twice the spin
    (`tentorch/symmetry/cat/su2/su2.cpp`, `su2::coefficients`): X[k1, k2, k3] =
    <j1 m1; j2 m2 | j3 m3> with m = 2k - j.  The missing state at each level is the part of the
    highest-weight state orthogonal to the lowered ones, with a positive m1 = j1 component.
    '''
    lo, hi, n1 = abs(j1 - j2), j1 + j2, j1 + 1
    out = {hi: np.zeros((j1 + 1, j2 + 1, hi + 1))}
    out[hi][j1, j2, hi] = 1.0
    previous = {hi: np.zeros(n1)}
    previous[hi][j1] = 1.0

    def lower(vec, MM):
        res = np.zeros(n1)
        for k1 in range(n1):
            m1 = 2 * k1 - j1
            m2 = MM + 2 - m1
            if abs(m2) > j2:
                continue
            if abs(m2 - 2) <= j2:
                res[k1] += vec[k1] * math.sqrt((j2 + m2) * (j2 - m2 + 2)) / 2
            if k1 > 0 and abs(m1 - 2) <= j1:
                res[k1 - 1] += vec[k1] * math.sqrt((j1 + m1) * (j1 - m1 + 2)) / 2
        return res

    for MM in range(hi - 2, -hi - 1, -2):
        level = {}
        for JJ in sorted(previous, reverse=True):
            if JJ < abs(MM):
                continue
            scale = math.sqrt((JJ + MM + 2) * (JJ - MM)) / 2
            level[JJ] = lower(previous[JJ], MM) / scale
        if MM >= lo and MM >= 0:
            vec = np.zeros(n1)
            vec[j1] = 1.0
            for JJ in sorted(level, reverse=True):
                vec = vec - np.dot(level[JJ], vec) * level[JJ]
            norm = math.sqrt(float(np.dot(vec, vec)))
            if norm <= 1e-12:
                raise RuntimeError('degenerate orthogonalisation')
            level[MM] = vec / norm
        previous = level
        for JJ, vec in level.items():
            if JJ not in out:
                out[JJ] = np.zeros((j1 + 1, j2 + 1, JJ + 1))
            for k1 in range(n1):
                m1 = 2 * k1 - j1
                m2 = MM - m1
                if abs(m2) <= j2 and vec[k1] != 0.0:
                    out[JJ][k1, (m2 + j2) // 2, (MM + JJ) // 2] += vec[k1]
    return out


class SU2Symmetry(Group):
    '''`SU2Symmetry` (`tentorch/symmetry/cat/su2/su2.cpp`): sectors are 2j.'''
    name = 'SU(2)'

    def __init__(self) -> None:
        super().__init__()
        self._tables: Dict[Tuple[int, int], Dict[int, np.ndarray]] = {}

    def compute_outcomes(self, a, b):
        return [((jj,), 1) for jj in range(abs(a[0] - b[0]), a[0] + b[0] + 1, 2)]

    def dual(self, a):
        return a

    def q_dim(self, a):
        return float(a[0] + 1)

    def frobenius_schur_sign(self, a):
        return 1.0 if a[0] % 2 == 0 else -1.0

    def compute_cg(self, a, b, c, copy):
        if copy != 0:
            raise RuntimeError('SU(2) couples each pair once')
        key = (a[0], b[0])
        if key not in self._tables:
            self._tables[key] = su2_coefficients(a[0], b[0])
        return self._tables[key][c[0]]


class FibonacciCategory(Category):
    '''`FibonacciCategory` (`tentorch/symmetry/cat/fibonacci/fibonacci.cpp:20-79`).'''
    name = 'FibonacciCategory'

    def compute_outcomes(self, a, b):
        if a[0] == 0:
            return [(b, 1)]
        if b[0] == 0:
            return [(a, 1)]
        return [((0,), 1), ((1,), 1)]

    def compute_F(self, a, b, c, d):
        rows = self.f_rows(a, b, c, d)
        cols = self.f_columns(a, b, c, d)
        value = np.zeros((len(rows), len(cols)), dtype=np.complex128)
        if not rows or not cols:
            return value
        if a == b == c == d == (1,):
            phi = (1 + math.sqrt(5)) / 2
            golden = [[1 / phi, 1 / math.sqrt(phi)], [1 / math.sqrt(phi), -1 / phi]]
            for i, r in enumerate(rows):
                for j, col in enumerate(cols):
                    value[i, j] = golden[r[0][0]][col[0][0]]
            return value
        value[:] = 1.0
        return value

    def compute_R(self, b, c):
        rows = self.r_rows(b, c)
        value = np.zeros((len(rows), len(rows)), dtype=np.complex128)
        for i, (e, _) in enumerate(rows):
            phase = 1.0 + 0j
            if b == (1,) and c == (1,):
                phase = np.exp(-4j * math.pi / 5) if e == (0,) else np.exp(3j * math.pi / 5)
            value[i, i] = phase
        return value

    def q_dim(self, a):
        return (1 + math.sqrt(5)) / 2 if a[0] == 1 else 1.0

    def dual(self, a):
        return a


no_symmetry = NoSymmetry()
u1 = U1Symmetry()
fermion_parity = FermionParity()
su2 = SU2Symmetry()
fibonacci = FibonacciCategory()
_zn: Dict[int, ZNSymmetry] = {}


def z_n(modulus: int) -> ZNSymmetry:
    if modulus not in _zn:
        _zn[modulus] = ZNSymmetry(modulus)
    return _zn[modulus]


class Leg:
    '''
    `Leg` (`tentorch/symmetry/tree/tree.cpp:90-140`): sectors sorted and unique, multiplicities
    positive.
    '''

    def __init__(self, cat: Category, sectors: Sequence[Sector], multiplicities: Sequence[int]):
        sectors = [tuple(int(v) for v in s) for s in sectors]
        multiplicities = [int(m) for m in multiplicities]
        if len(sectors) != len(multiplicities):
            raise ValueError(f'A leg needs one multiplicity per sector, got {len(multiplicities)} '
                             f'for {len(sectors)} sectors.')
        for m in multiplicities:
            if m <= 0:
                raise ValueError(f'A block multiplicity must be positive, got {m}.')
        order = sorted(range(len(sectors)), key=lambda i: sectors[i])
        self.cat = cat
        self.sectors: List[Sector] = [sectors[i] for i in order]
        self.multiplicities: List[int] = [multiplicities[i] for i in order]
        for i in range(1, len(self.sectors)):
            if self.sectors[i] == self.sectors[i - 1]:
                raise ValueError(f'A leg cannot hold the sector {render(self.sectors[i])} twice.')
        self._index = {s: i for i, s in enumerate(self.sectors)}

    def nblocks(self) -> int:
        return len(self.sectors)

    def sector(self, block: int) -> Sector:
        return self.sectors[block]

    def multiplicity(self, block: int) -> int:
        return self.multiplicities[block]

    def multiplets(self) -> int:
        return sum(self.multiplicities)

    def block_of(self, sector: Sector) -> int:
        '''`block_of` (`tentorch/tensors/layout.cpp:8-16`).'''
        found = self._index.get(sector)
        if found is None:
            raise ValueError(f'a leg of {self.nblocks()} blocks carries none in sector {render(sector)}')
        return found

    def mult_of(self, sector: Sector) -> int:
        return self.multiplicities[self.block_of(sector)]

    def fuse(self, other: 'Leg') -> 'Leg':
        if self.cat is not other.cat:
            raise ValueError('Cannot fuse legs of different categories.')
        fused: Dict[Sector, int] = {}
        for a, ma in zip(self.sectors, self.multiplicities):
            for b, mb in zip(other.sectors, other.multiplicities):
                for c, n in self.cat.outcomes(a, b):
                    fused[c] = fused.get(c, 0) + ma * mb * n
        keys = sorted(fused)
        return Leg(self.cat, keys, [fused[k] for k in keys])

    def conj(self) -> 'Leg':
        return Leg(self.cat, [self.cat.dual(s) for s in self.sectors], self.multiplicities)

    def capped(self, max_multiplets: int) -> 'Leg':
        if max_multiplets <= 0:
            raise ValueError(f'A cap must keep at least one multiplet, got {max_multiplets}.')
        sectors, mults, kept = [], [], 0
        for s, m in zip(self.sectors, self.multiplicities):
            if kept >= max_multiplets:
                break
            take = min(m, max_multiplets - kept)
            sectors.append(s)
            mults.append(take)
            kept += take
        return Leg(self.cat, sectors, mults)

    def __eq__(self, other) -> bool:
        return (isinstance(other, Leg) and self.cat is other.cat and self.sectors == other.sectors
                and self.multiplicities == other.multiplicities)

    def __hash__(self) -> int:
        return hash((self.cat.name, tuple(self.sectors), tuple(self.multiplicities)))

    def key(self) -> tuple:
        return (id(self.cat), tuple(self.sectors), tuple(self.multiplicities))

    def __repr__(self) -> str:
        return f'Leg({self.cat.name}, {self.sectors}, {self.multiplicities})'


class FusionTree:
    '''`FusionTree` (`tentorch/symmetry/tree/tree.h:134-155`): left-nested.'''
    __slots__ = ('sectors', 'inner', 'copies', 'coupled')

    def __init__(self, sectors, inner, copies, coupled):
        self.sectors = tuple(sectors)
        self.inner = tuple(inner)
        self.copies = tuple(copies)
        self.coupled = coupled

    def key(self):
        return (self.sectors, self.inner, self.copies, self.coupled)

    def size(self):
        return len(self.sectors)

    def __eq__(self, other):
        return self.key() == other.key()

    def __hash__(self):
        return hash(self.key())

    def __repr__(self):
        return f'FusionTree({self.sectors}, {self.inner}, {self.copies}, {self.coupled})'


def reachable_charges(cat: Category, legs: Sequence[Leg]) -> List[Sector]:
    '''`reachable_charges` (`tree.cpp:375-398`): insertion order.'''
    charges = [cat.trivial()]
    for leg in legs:
        nxt, seen = [], set()
        for charge in charges:
            for s in leg.sectors:
                for c, _ in cat.outcomes(charge, s):
                    if c not in seen:
                        seen.add(c)
                        nxt.append(c)
        charges = nxt
    return charges


def fusion_trees(cat: Category, legs: Sequence[Leg], coupled: Sector) -> List[FusionTree]:
    '''`fusion_trees` (`tree.cpp:400-471`): depth-first, leg blocks in sorted order.'''
    count = len(legs)
    if count == 0:
        return [FusionTree((), (), (), coupled)] if coupled == cat.trivial() else []
    out: List[FusionTree] = []
    sectors = [None] * count
    inner = [None] * max(count - 2, 0)
    copies = [0] * max(count - 1, 0)

    def walk(leg: int, previous: Sector) -> None:
        for s in legs[leg].sectors:
            sectors[leg] = s
            if leg == 0:
                if count == 1:
                    if s == coupled:
                        out.append(FusionTree(sectors, (), (), coupled))
                else:
                    walk(1, s)
                continue
            for c, mult in cat.outcomes(previous, s):
                if leg + 1 == count and c != coupled:
                    continue
                if leg + 1 < count:
                    inner[leg - 1] = c
                for copy in range(mult):
                    copies[leg - 1] = copy
                    if leg + 1 == count:
                        out.append(FusionTree(sectors, inner, copies, coupled))
                    else:
                        walk(leg + 1, c)

    walk(0, cat.trivial())
    return out


def trimmed(cat: Category, tree: FusionTree) -> FusionTree:
    '''`trimmed` (`tree.cpp:473-484`).'''
    n = tree.size()
    coupled = cat.trivial() if n <= 1 else (tree.sectors[0] if n == 2 else tree.inner[n - 3])
    return FusionTree(tree.sectors[:-1], tree.inner[:-1], tree.copies[:-1], coupled)


class SideLayout:
    '''`SideLayout` (`tentorch/tensors/layout.h:28-35`, `layout.cpp:18-32`).'''
    __slots__ = ('trees', 'offsets', 'index')

    def __init__(self, trees, offsets):
        self.trees = trees
        self.offsets = offsets
        self.index = {t.key(): i for i, t in enumerate(trees)}

    def rows(self):
        return self.offsets[-1]


_layouts: Dict[tuple, SideLayout] = {}


def layout_of(cat: Category, legs: Sequence[Leg], root: Sector) -> SideLayout:
    '''`layout_of` (`layout.cpp:18-32`): one row per multiplicity combination, last leg fastest.'''
    key = (id(cat), tuple(l.key() for l in legs), root)
    found = _layouts.get(key)
    if found is not None:
        return found
    trees = fusion_trees(cat, legs, root)
    offsets = [0]
    for tree in trees:
        rows = 1
        for leg, s in zip(legs, tree.sectors):
            rows *= leg.mult_of(s)
        offsets.append(offsets[-1] + rows)
    found = SideLayout(trees, offsets)
    _layouts[key] = found
    return found


def block_keys(cat: Category, codomain: Sequence[Leg], domain: Sequence[Leg], qtotal: Sector):
    '''`block_keys` (`layout.cpp:34-47`): (r, s, copy) with N(r, dual(s), q) > 0.'''
    keys = []
    for r in reachable_charges(cat, codomain):
        for s in reachable_charges(cat, domain):
            for copy in range(cat.n_symbol(r, cat.dual(s), qtotal)):
                keys.append((r, s, copy))
    return keys
