'''
The Heisenberg chain as a matrix product operator (`tentorch/models/heisenberg.py`) and the model
container (`tentorch/models/model.py`).
'''

from dataclasses import dataclass
from typing import Optional

from .networks import MPO, entered
from .symmetry import Leg, is_abelian_group, is_no_symmetry, u1

# The automaton of `heisenberg.py:15-37`: 0 nothing written, 1 S^z, 2 S^+, 3 S^-, 4 closed.
START, CLOSED = 0, 4
_STATE_OF_CHARGED = {(0,): {0: START, 1: 1, 2: CLOSED}, (2,): {0: 2}, (-2,): {0: 3}}
_OPERATOR = {
    (START, START): lambda out, into: 1.0 if out == into else 0.0,
    (CLOSED, CLOSED): lambda out, into: 1.0 if out == into else 0.0,
    (START, 1): lambda out, into: 0.5 if out == into == 1 else (-0.5 if out == into else 0.0),
    (1, CLOSED): lambda out, into: 0.5 if out == into == 1 else (-0.5 if out == into else 0.0),
    (START, 2): lambda out, into: 1.0 if out == into + 2 else 0.0,
    (2, CLOSED): lambda out, into: 0.5 if out == into - 2 else 0.0,
    (START, 3): lambda out, into: 1.0 if out == into - 2 else 0.0,
    (3, CLOSED): lambda out, into: 0.5 if out == into + 2 else 0.0,
}


@dataclass
class Model:
    '''`Model` (`tentorch/models/model.py:7-34`).'''
    length: int
    coupling: complex = 1.0
    cat: object = u1
    physical: Optional[object] = None
    sector: Optional[tuple] = None
    state: Optional[object] = None
    energy: Optional[float] = None

    def __post_init__(self):
        if self.length < 1:
            raise ValueError(f'a chain needs at least one site, not {self.length}')
        if self.sector is None:
            self.sector = self.cat.trivial()


def heisenberg(cat, length: int, coupling: float = 1.0, dtype=None, device=None) -> MPO:
    '''`heisenberg` (`tentorch/models/heisenberg.py:58-115`).'''
    if not is_abelian_group(cat):
        raise ValueError(f'the rows of {cat.name} would need the recoupling to dress them')
    if is_no_symmetry(cat):
        phys, bond = Leg(cat, [(0,)], (2,)), Leg(cat, [(0,)], (5,))
        state_of = lambda sector, copy: copy
        spin_of = lambda sector, copy: 1 - 2 * copy
    else:
        phys = Leg(cat, [(1,), (-1,)], (1, 1))
        bond = Leg(cat, [(0,), (2,), (-2,)], (3, 1, 1))
        state_of = lambda sector, copy: _STATE_OF_CHARGED[tuple(sector)][copy]
        spin_of = lambda sector, copy: sector[0]
    edge = Leg(cat, [cat.trivial()], (1,))
    coupling = complex(coupling).real

    def site(w_left, w_right, first, last):
        def entry(t_out, c_out, t_in, c_in):
            before = START if first else state_of(t_out.sectors[0], c_out[0])
            after = CLOSED if last else state_of(t_in.sectors[1], c_in[1])
            write = _OPERATOR.get((before, after))
            if write is None:
                return 0.0
            weight = coupling if before not in (START, CLOSED) and after == CLOSED else 1.0
            return weight * write(spin_of(t_out.sectors[1], c_out[1]), spin_of(t_in.sectors[0], c_in[0]))
        return entered(cat, [w_left, phys], [phys, w_right], entry, dtype=dtype, device=device)

    sites = [site(edge if at == 0 else bond, edge if at == length - 1 else bond, at == 0, at == length - 1)
             for at in range(length)]
    return MPO(sites, phys, cat)
