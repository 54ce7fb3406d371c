"""IM model description flattened for the genealogy-prior kernels.

Periods k = 0..npops-1 are the intervals between split times; period k
holds the populations in ``plist[k]`` (sampled populations first, then
ancestral ones as they appear).  A population ends at the period where it
merges (``pope``, -1 for the root population) into ``popdown``.  The split
closing period k merges ``join[k] = (a, b)`` into population npops + k,
the reference's droppops/addpop (build_poptree.cpp).

Parameters follow the reference's default model (setup_iparams,
initialize.cpp:202-760): one population-size parameter per population,
summed over every period the population is present in; one migration
parameter per ordered pair of populations that coexist in a period before
the last, summed over every such period.  Priors are uniform on
[0, qmax] and [0, mmax].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .sim import PopTree


@dataclass
class IMModel:
    npops: int
    nper: int
    plist: np.ndarray      # [nper, npops] int32 (-1 padded)
    npresent: np.ndarray   # [nper]
    popdown: np.ndarray    # [2npops-1]
    pope: np.ndarray       # [2npops-1]
    join: np.ndarray       # [nper, 3] (a, b, new); last row unused
    theta_wp: list         # per theta param: [(k, i)]
    mig_wp: list           # per migration param: [(k, i, j)]
    qmax: np.ndarray
    mmax: np.ndarray

    @property
    def npsize(self):
        return len(self.theta_wp)

    @property
    def nmig(self):
        return len(self.mig_wp)


def build(pt: PopTree, qmax=20.0, mmax=5.0, migration=True) -> IMModel:
    """migration=False: no migration parameters (the reference's model
    without migration, -j option)."""
    P = pt.npops
    nper = P
    plist = -np.ones((nper, P), np.int32)
    npresent = np.zeros(nper, np.int32)
    for k in range(nper):
        act = pt.active(k)
        # keep earlier positions stable: survivors first, new population last
        if k > 0:
            prev = [p for p in plist[k - 1][: npresent[k - 1]] if p in act]
            act = prev + [p for p in act if p not in prev]
        plist[k, : len(act)] = act
        npresent[k] = len(act)
    ntree = 2 * P - 1
    popdown = -np.ones(ntree, np.int32)
    pope = -np.ones(ntree, np.int32)
    join = -np.ones((nper, 3), np.int32)
    for k, (a, b) in enumerate(pt.joins):
        new = P + k
        popdown[a] = popdown[b] = new
        pope[a] = pope[b] = k + 1
        join[k] = (a, b, new)
    theta_wp = []
    for p in range(ntree):
        theta_wp.append([(k, int(np.where(plist[k] == p)[0][0]))
                         for k in range(nper) if p in plist[k][: npresent[k]]])
    mig_wp, pairs = [], {}
    for k in range(nper - 1 if migration else 0):
        pres = list(plist[k][: npresent[k]])
        for i, p in enumerate(pres):
            for j, q in enumerate(pres):
                if i == j:
                    continue
                key = (int(p), int(q))
                if key not in pairs:
                    pairs[key] = len(mig_wp)
                    mig_wp.append([])
                mig_wp[pairs[key]].append((k, i, j))
    qm = np.broadcast_to(np.asarray(qmax, float), (len(theta_wp),)).copy()
    mm = np.broadcast_to(np.asarray(mmax, float), (len(mig_wp),)).copy()
    return IMModel(P, nper, plist, npresent, popdown, pope, join, theta_wp, mig_wp, qm, mm)


def mig_arrays(g):
    """Migration events of genealogy g sorted by edge, then time."""
    e, t, to = [], [], []
    for edge, lst in enumerate(g.mig):
        for (mt, mp) in sorted(lst):
            e.append(edge)
            t.append(mt)
            to.append(mp)
    return (np.array(e, np.int32), np.array(t, np.float64), np.array(to, np.int32))
