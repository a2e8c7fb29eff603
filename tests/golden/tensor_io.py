'''
Block-sparse tensors in the reference's `to_tuple` layout (`tentorch/_ext/bindings.cpp:311-323`:
codomain legs, domain legs, total charge, {block key: block}) stored as an .npz of plain arrays: for
a tensor named T, `T.legs` holds every leg as JSON ([side, sectors, multiplicities] per leg), `T.keys`
the block keys as JSON ([codomain root, domain root, copy] per block, in key order) and `T.block<i>`
the i-th block.  Written and read without pickling.
'''

import json

import numpy as np


def pack(name, codomain, domain, qtotal, blocks):
    '''Arrays for one tensor: legs as (sectors, multiplicities) pairs, blocks keyed by
    (codomain root, domain root, copy).'''
    legs = [['codomain', [list(s) for s in sectors], list(mults)] for sectors, mults in codomain] + \
           [['domain', [list(s) for s in sectors], list(mults)] for sectors, mults in domain]
    keys = list(blocks)
    out = {f'{name}.legs': np.array(json.dumps({'legs': legs, 'qtotal': list(qtotal)})),
           f'{name}.keys': np.array(json.dumps([[list(r), list(s), int(c)] for r, s, c in keys]))}
    for i, key in enumerate(keys):
        out[f'{name}.block{i}'] = np.ascontiguousarray(blocks[key])
    return out


def unpack(data, name):
    '''(codomain, domain, qtotal, {key: block}) of one tensor; a leg is (sectors, multiplicities).'''
    meta = json.loads(str(data[f'{name}.legs']))
    codomain = [([tuple(s) for s in sectors], mults) for side, sectors, mults in meta['legs'] if side == 'codomain']
    domain = [([tuple(s) for s in sectors], mults) for side, sectors, mults in meta['legs'] if side == 'domain']
    keys = [(tuple(r), tuple(s), c) for r, s, c in json.loads(str(data[f'{name}.keys']))]
    return codomain, domain, tuple(meta['qtotal']), {key: data[f'{name}.block{i}'] for i, key in enumerate(keys)}


def load(path):
    return np.load(path, allow_pickle=False)
