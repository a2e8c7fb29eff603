'''
The fusion categories and legs of the MI355X backend, as the host engine implements them
(`tentorch_amd/csrc/engine/category.cpp`, `tree.cpp`), under the reference's names
(`tentorch/symmetry/__init__.py`).  A sector is a tuple of ints.
'''

from collections import namedtuple

from ._native import engine

FusionCategory = engine.FusionCategory
Leg = engine.Leg
Sector = tuple

no_symmetry = engine.no_symmetry()
u1 = engine.u1()
fermion_parity = engine.fermion_parity()
su2 = engine.su2()
fibonacci = engine.fibonacci()


def z_n(modulus: int):
    '''The cyclic group Z_N (`tentorch/symmetry/cat/abelian/abelian.cpp:74-81`).'''
    return engine.z_n(modulus)


class FusionTree(namedtuple('FusionTree', ['sectors', 'inner', 'copies', 'coupled'])):
    '''A left-nested fusion tree (`tentorch/symmetry/tree/tree.h:134-155`).'''

    def size(self) -> int:
        return len(self.sectors)


def fusion_trees(cat, legs, coupled):
    '''`fusion_trees` (`tentorch/symmetry/tree/tree.cpp:400-471`).'''
    return [FusionTree(*t) for t in engine.fusion_trees(cat, list(legs), tuple(coupled))]


def reachable_charges(cat, legs):
    '''`reachable_charges` (`tree.cpp:375-398`).'''
    return engine.reachable_charges(cat, list(legs))


def is_abelian_group(cat) -> bool:
    '''Whether the category is one of the abelian groups (`AbelianGroup` in the reference).'''
    return cat.abelian and cat.name != 'FermionParity'


def is_no_symmetry(cat) -> bool:
    return cat.name == 'NoSymmetry'
