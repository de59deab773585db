"""Jastrow factor from a TurboRVB fort.10 text file (the reference's fixtures
tests/trexio_example_files/turborvb_WF_*.txt), built the way the reference's converter
tests/trexio_example_files/read_jastrow_factor_from_turbo_wf.py builds its Jastrow_data:

* two-body parameter (Pade) and, for Jastrow type -15, the one-body parameter (exp form);
* J3 basis: one normalised real-spherical Gaussian shell per Jastrow shell (makefun m order
  p: +1, -1, 0; d: 0, +2, -2, -1, +1; higher l: 0, +1, -1, ...), coefficient 1;
* J matrix: the constant orbital is the largest index; its column (times N_el - 1) is the
  j1 vector, every other non-zero (row, col) is mirrored into the symmetric J3 block.
"""

import numpy as np

from jqmc_amd.atomic_orbital import AOs_sphe_data
from jqmc_amd.jastrow_factor import Jastrow_data, Jastrow_one_body_data, Jastrow_three_body_data, Jastrow_two_body_data

_M_ORDER = {0: [0], 1: [+1, -1, 0], 2: [0, +2, -2, -1, +1]}


def _tokens_after(lines, marker):
    for i, line in enumerate(lines):
        if marker in line:
            toks = []
            for rest in lines[i + 1 :]:
                if rest.lstrip().startswith("#"):
                    break
                toks += rest.split()
            return toks
    raise KeyError(marker)


def _num(t):
    return float(t.replace("D", "E").replace("d", "e"))


def read_jastrow(path, structure_data, twobody_first: bool = True):
    lines = open(path).read().splitlines()
    nelup, nel, nion = (int(x) for x in _tokens_after(lines, "# Nelup")[:3])
    n_shell_det, n_shell_jas = (int(x) for x in _tokens_after(lines, "# Shell Det.")[:2])
    jas_type = int(_tokens_after(lines, "# Jas 2body")[0])
    n_detmat, n_jasmat = (int(x) for x in _tokens_after(lines, "# Det mat.")[:2])
    ion = [_num(t) for t in _tokens_after(lines, "# Ion coordinates")[: 5 * nion]]
    zval = [ion[5 * i] for i in range(nion)]
    zat = [ion[5 * i + 1] for i in range(nion)]
    core = [z - v for z, v in zip(zat, zval)]
    tb = _tokens_after(lines, "Parameters Jastrow two body")
    params = [_num(t) for t in tb[1 : 1 + int(tb[0])]]
    toks = _tokens_after(lines, "Parameters atomic Jastrow wf")
    p = 0
    shells = []
    for _ in range(n_shell_jas):
        mult, npar, typ = int(toks[p]), int(toks[p + 1]), int(toks[p + 2])
        nuc = int(toks[p + 3]) - 1
        vals = [_num(t) for t in toks[p + 4 : p + 4 + npar]]
        p += 4 + npar
        shells.append((mult, typ, nuc, vals))
    jm = _tokens_after(lines, "Nonzero values of  jasmat")
    entries = [(int(jm[3 * k]), int(jm[3 * k + 1]), _num(jm[3 * k + 2])) for k in range(n_jasmat)]

    nuc_idx, orb_idx, exps, coefs, angs, mags = [], [], [], [], [], []
    ao = 0
    for mult, typ, nuc, vals in shells:
        if typ == 200:  # constant orbital, carried by the j1 column
            continue
        l = (mult - 1) // 2
        ms = _M_ORDER.get(l, [0] + [i * (-1) ** j for i in range(1, l + 1) for j in range(2)])
        for m in ms:
            for z in vals:
                exps.append(z)
                coefs.append(1.0)
                orb_idx.append(ao)
            nuc_idx.append(nuc)
            angs.append(l)
            mags.append(m)
            ao += 1
    aos = AOs_sphe_data(structure_data, nuc_idx, ao, len(exps), orb_idx, np.array(exps), np.array(coefs), angs, mags)
    const = max(r for r, _, _ in entries)
    j1 = np.zeros(ao)
    j3 = np.zeros((ao, ao))
    for r, c, v in entries:
        if c == const and r != const:
            j1[r - 1] = v * (nel - 1)
        elif r != const and c != const:
            j3[r - 1, c - 1] = v
            j3[c - 1, r - 1] = v
    if jas_type == -5:
        two, one = params[0], None
    elif jas_type == -15:
        two, one = (params[0], params[1]) if twobody_first else (params[1], params[0])
    else:
        raise NotImplementedError(f"Jastrow type {jas_type}")
    j1d = None
    if one is not None:
        j1d = Jastrow_one_body_data(jastrow_1b_param=one, jastrow_1b_type="exp", structure_data=structure_data,
                                    core_electrons=core)
    return Jastrow_data(
        jastrow_one_body_data=j1d,
        jastrow_two_body_data=Jastrow_two_body_data(jastrow_2b_param=two, jastrow_2b_type="pade"),
        jastrow_three_body_data=Jastrow_three_body_data(orb_data=aos, j_matrix=np.column_stack((j3, j1))),
    )
