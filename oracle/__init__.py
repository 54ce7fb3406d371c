"""ctypes front end of the oracle (TEST INFRASTRUCTURE ONLY).

The oracle is the CPU restatement of the reference hot path in
``oracle/ima_oracle.c``.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this module: it is the checker,
never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libima_oracle.so")
REJECT_IS = -1000000000.0


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def _load():
    if not os.path.exists(_SO):
        build()
    lib = C.CDLL(_SO)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    fp = C.POINTER(C.c_float)
    lib.ora_getstandfactor.restype = C.c_double
    lib.ora_getstandfactor.argtypes = [dp, C.c_double]
    lib.ora_likelihoodHKY.restype = C.c_double
    lib.ora_likelihoodHKY.argtypes = [C.c_void_p, C.c_int, ip, ip, C.c_int, dp,
                                      C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_void_p]
    for f in ("ora_copyfraclike", "ora_storescalefactors", "ora_restorescalefactors"):
        getattr(lib, f).restype = None
        getattr(lib, f).argtypes = [C.c_void_p, C.c_void_p]
    lib.ora_labelgtree.restype = None
    lib.ora_labelgtree.argtypes = [C.c_void_p, ip, C.c_int]
    lib.ora_calc_sumlogk.restype = C.c_int
    lib.ora_calc_sumlogk.argtypes = [C.c_void_p, C.c_int, ip, dp]
    lib.ora_likelihoodIS.restype = C.c_double
    lib.ora_likelihoodIS.argtypes = [C.c_void_p, C.c_int, ip, C.c_double,
                                     C.c_double, C.c_double, ip, ip]
    lib.ora_tree_length.restype = C.c_double
    lib.ora_tree_length.argtypes = [C.c_void_p]
    for f in ("ora_bessi0", "ora_bessi1"):
        getattr(lib, f).restype = C.c_double
        getattr(lib, f).argtypes = [C.c_double]
    lib.ora_bessi.restype = C.c_double
    lib.ora_bessi.argtypes = [C.c_int, C.c_double]
    lib.ora_likelihoodSW.restype = C.c_double
    lib.ora_likelihoodSW.argtypes = [C.c_void_p, ip, C.c_double, C.c_double, dp]
    lib.ora_finishSWupdateA.restype = C.c_double
    lib.ora_finishSWupdateA.argtypes = [C.c_void_p, ip, C.c_void_p, C.c_int, dp, ip, dp]
    lib.ora_swapweight.restype = C.c_double
    lib.ora_swapweight.argtypes = [C.c_double] * 4
    lib.ora_setheat_geometric.restype = C.c_double
    lib.ora_setheat_geometric.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
    lib.ora_setheat_linear.restype = C.c_double
    lib.ora_setheat_linear.argtypes = [C.c_int, C.c_double]
    lib.ora_eexp.restype = None
    lib.ora_eexp.argtypes = [C.c_double, dp, ip]
    lib.ora_jointp.restype = C.c_double
    lib.ora_jointp.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.c_int, C.c_int, dp, C.c_int, dp]
    lib.ora_jointp_exact.restype = C.c_double
    lib.ora_jointp_exact.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_int, C.c_int, dp]
    lib.ora_margincalc.restype = C.c_double
    lib.ora_margincalc.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.c_double, C.c_int, C.c_int,
                                   C.c_int, C.c_double]
    lib.ora_popmig.restype = C.c_double
    lib.ora_popmig.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_double, C.c_double, C.c_int, C.c_int, C.c_double]
    lib.ora_marginpopmig.restype = C.c_double
    lib.ora_marginpopmig.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                     C.c_double]
    lib.ora_popmig_peak.restype = C.c_double
    lib.ora_popmig_peak.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, dp]
    lib.ora_gtpops_one.restype = C.c_double
    lib.ora_gtpops_one.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
    lib.ora_gtmig_one.restype = C.c_double
    lib.ora_gtmig_one.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
    lib.ora_greater_than.restype = C.c_double
    lib.ora_greater_than.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_double]
    lib.ora_mc3_bench.restype = C.c_long
    lib.ora_mc3_bench.argtypes = ([C.c_int, C.c_int] + [ip] * 7 + [dp, dp, dp, ip, C.c_int]
                                  + [ip] * 5 + [C.c_int, ip, ip, C.c_int, ip, ip, dp, dp,
                                                C.c_double] + [ip] * 4
                                  + [dp, ip, ip, ip, dp, ip, dp, dp, C.c_double, C.c_uint64,
                                     C.c_double, C.c_long, C.POINTER(C.c_long), C.c_int,
                                     C.c_int, C.c_int, ip, C.c_int, ip])
    lib.ora_getmprob_nw.restype = C.c_double
    lib.ora_getmprob_nw.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double] + [C.c_int] * 5
    lib.ora_poisson_cond.restype = C.c_int
    lib.ora_poisson_cond.argtypes = [C.c_double, C.c_int, C.c_uint64, C.c_uint64]
    lib.ora_update_mig_tnw.restype = C.c_int
    lib.ora_update_mig_tnw.argtypes = ([C.c_int, ip, ip, ip, dp, C.c_int, ip, ip, C.c_int, ip, dp,
                                        ip, C.c_int] + [ip] * 5
                                       + [dp, C.c_int, C.c_double, C.c_uint64, C.c_uint64, dp, ip])
    return lib


lib = _load()


class _Tree(C.Structure):
    _fields_ = [("n", C.c_int), ("up0", C.POINTER(C.c_int)),
                ("up1", C.POINTER(C.c_int)), ("down", C.POINTER(C.c_int)),
                ("time", C.POINTER(C.c_double)), ("root", C.c_int)]


class _HKYState(C.Structure):
    _fields_ = [("n", C.c_int), ("S", C.c_int),
                ("frac", C.POINTER(C.c_double)), ("newfrac", C.POINTER(C.c_double)),
                ("scale", C.POINTER(C.c_double)), ("oldscale", C.POINTER(C.c_double))]


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Tree:
    """Keeps numpy buffers alive for the C struct."""

    def __init__(self, g):
        self.n = int(g.n)
        self.up0 = np.ascontiguousarray(g.up0, np.int32)
        self.up1 = np.ascontiguousarray(g.up1, np.int32)
        self.down = np.ascontiguousarray(g.down, np.int32)
        self.time = np.ascontiguousarray(g.time, np.float64)
        self.s = _Tree(self.n, _ip(self.up0), _ip(self.up1), _ip(self.down),
                       _dp(self.time), int(g.root))

    @property
    def ref(self):
        return C.byref(self.s)


class HKYState:
    def __init__(self, n, S):
        self.frac = np.zeros((n - 1) * S * 4)
        self.newfrac = np.zeros((n - 1) * S * 4)
        self.scale = np.zeros((n - 1) * S)
        self.oldscale = np.zeros((n - 1) * S)
        self.s = _HKYState(n, S, _dp(self.frac), _dp(self.newfrac),
                           _dp(self.scale), _dp(self.oldscale))

    @property
    def ref(self):
        return C.byref(self.s)


def getstandfactor(pi, kappa):
    pi = np.ascontiguousarray(pi, np.float64)
    return lib.ora_getstandfactor(_dp(pi), kappa)


def likelihood_hky(g, locus, u, kappa, e=(-1, -1, -1, -1), state=None,
                   commit=True):
    """likelihoodHKY on genealogy g; a full evaluation commits its partials
    (copyfraclike + storescalefactors) when commit is set."""
    t = Tree(g)
    S = locus.seq.shape[1]
    if state is None:
        state = HKYState(g.n, S)
    seq = np.ascontiguousarray(locus.seq, np.int32)
    mult = np.ascontiguousarray(locus.mult, np.int32)
    pi = np.ascontiguousarray(locus.pi, np.float64)
    v = lib.ora_likelihoodHKY(t.ref, S, _ip(seq), _ip(mult), int(locus.totsites),
                              _dp(pi), float(u), float(kappa), *[int(x) for x in e],
                              state.ref)
    if commit:
        lib.ora_copyfraclike(t.ref, state.ref)
        lib.ora_storescalefactors(t.ref, state.ref)
    return v, state


def hky_commit(g, state, accept):
    t = Tree(g)
    if accept:
        lib.ora_copyfraclike(t.ref, state.ref)
        lib.ora_storescalefactors(t.ref, state.ref)
    else:
        lib.ora_restorescalefactors(t.ref, state.ref)


def labelgtree_site(g, col):
    """Run the reference's site labelling; returns labels of all edges."""
    t = Tree(g)
    mut = np.full(2 * g.n - 1, -1, np.int32)
    mut[: g.n] = col
    for j in range(g.n):
        lib.ora_labelgtree(t.ref, _ip(mut), j)
    return mut


def calc_sumlogk(g, locus):
    t = Tree(g)
    seq = np.ascontiguousarray(locus.seq, np.int32)
    out = np.zeros(1)
    rc = lib.ora_calc_sumlogk(t.ref, seq.shape[1], _ip(seq), _dp(out))
    return rc, float(out[0])


def tree_length(g):
    return lib.ora_tree_length(Tree(g).ref)


def likelihood_is(g, locus, u, sumlogk, length=None):
    t = Tree(g)
    seq = np.ascontiguousarray(locus.seq, np.int32)
    S = seq.shape[1]
    if length is None:
        length = lib.ora_tree_length(t.ref)
    st = np.zeros(1, np.int32)
    nodes = np.full(max(S, 1), -1, np.int32)
    v = lib.ora_likelihoodIS(t.ref, S, _ip(seq), float(length), float(u),
                             float(sumlogk), _ip(st), _ip(nodes))
    return v, int(st[0]), nodes[:S]


def bessi(n, x):
    return lib.ora_bessi(int(n), float(x))


def bessi0(x):
    return lib.ora_bessi0(float(x))


def bessi1(x):
    return lib.ora_bessi1(float(x))


def likelihood_sw(g, A, u, tr=1.0):
    t = Tree(g)
    A = np.ascontiguousarray(A, np.int32)
    d = np.zeros(2 * g.n - 1)
    v = lib.ora_likelihoodSW(t.ref, _ip(A), float(u), float(tr), _dp(d))
    return v, d


class _SWMove(C.Structure):
    _fields_ = [(f, C.c_int) for f in ("edge", "downedge", "sisedge", "newsisedge", "copyA0",
                                        "copyA1", "down_was_root", "hold_down_A", "minA",
                                        "maxA", "flip")] + \
               [("copy_dlikeA", C.c_double * 3), ("newsis_dlikeA", C.c_double),
                ("u", C.c_double), ("uni", C.c_double)]


def finish_sw_update_A(g, A, mv, calc=True):
    """finishSWupdateA on genealogy g (after the move) with allele states A;
    mv is a sim.SWMove.  Returns (dlike, Aterm, newA, edge dlikeA[4], A
    with the new allele)."""
    t = Tree(g)
    A = np.ascontiguousarray(A, np.int32).copy()
    m = _SWMove(mv.edge, mv.downedge, mv.sisedge, mv.newsisedge, mv.copyA0, mv.copyA1,
                int(mv.down_was_root), mv.hold_down_A, mv.minA, mv.maxA, int(mv.flip),
                (C.c_double * 3)(*mv.copy_dlikeA), mv.newsis_dlikeA, mv.u, mv.uni)
    at = np.zeros(1)
    na = np.zeros(1, np.int32)
    ed = np.zeros(4)
    v = lib.ora_finishSWupdateA(t.ref, _ip(A), C.byref(m), int(bool(calc)), _dp(at), _ip(na),
                                _dp(ed))
    return v, float(at[0]), int(na[0]), ed, A


def swapweight(bi, bj, si, sj):
    return lib.ora_swapweight(bi, bj, si, sj)


def setheat_geometric(i, nchains, h1, h2):
    return lib.ora_setheat_geometric(i, nchains, h1, h2)


def eexp(x):
    m = np.zeros(1)
    z = np.zeros(1, np.int32)
    lib.ora_eexp(float(x), _dp(m), _ip(z))
    return float(m[0]), int(z[0])


def jointp(table, npsize, nmig, npops, x, modeltype=0, prange=None, ess=False):
    table = np.ascontiguousarray(table, np.float32)
    G, K = table.shape
    x = np.ascontiguousarray(x, np.float64)
    r0, r1 = prange if prange is not None else (0, npsize + nmig)
    e = np.zeros(1)
    v = lib.ora_jointp(table.ctypes.data_as(C.POINTER(C.c_float)), G, K, npsize,
                       nmig, npops, modeltype, r0, r1, _dp(x), int(ess), _dp(e))
    return (v, float(e[0])) if ess else v


def jointp_exact(table, npsize, nmig, npops, x, modeltype=0, prange=None):
    table = np.ascontiguousarray(table, np.float32)
    G, K = table.shape
    x = np.ascontiguousarray(x, np.float64)
    r0, r1 = prange if prange is not None else (0, npsize + nmig)
    return lib.ora_jointp_exact(table.ctypes.data_as(C.POINTER(C.c_float)), G, K,
                                npsize, nmig, npops, modeltype, r0, r1, _dp(x))


def margincalc(table, npsize, nmig, x, pi, logi=1, yadjust=0.0, expo=False,
               mean=1.0):
    table = np.ascontiguousarray(table, np.float32)
    G, K = table.shape
    return lib.ora_margincalc(table.ctypes.data_as(C.POINTER(C.c_float)), G, K,
                              npsize, nmig, float(x), float(yadjust), int(pi),
                              int(logi), int(expo), float(mean))


def popmig(table, npsize, nmig, thetai, mi, x, qmax, mprior, expo=False, prob_or_like=0):
    """calc_popmig / calc_pop_expomig (popmig.cpp:9-172) at one x."""
    table = np.ascontiguousarray(table, np.float32)
    G, K = table.shape
    return lib.ora_popmig(table.ctypes.data_as(C.POINTER(C.c_float)), G, K, npsize, nmig,
                          thetai, mi, float(qmax), float(mprior), int(expo), int(prob_or_like),
                          float(x))


def marginpopmig(table, npsize, nmig, thetai, mi, x, qmax, mprior, expo=False, first=0,
                 last=None):
    """marginpopmig / marginpop_expomig (popmig.cpp:176-351) over rows [first, last)."""
    table = np.ascontiguousarray(table, np.float32)
    G, K = table.shape
    last = G if last is None else last
    return lib.ora_marginpopmig(table.ctypes.data_as(C.POINTER(C.c_float)), K, npsize, nmig,
                                thetai, mi, float(qmax), float(mprior), int(expo), int(first),
                                int(last), float(x))


def popmig_peak(table, npsize, nmig, thetai, mi, qmax, mprior, expo=False, first=0, last=None):
    """marginalopt_popmig (popmig.cpp:361-429) for one pair: (peak, density there)."""
    table = np.ascontiguousarray(table, np.float32)
    G, K = table.shape
    last = G if last is None else last
    pk = np.zeros(1)
    v = lib.ora_popmig_peak(table.ctypes.data_as(C.POINTER(C.c_float)), K, npsize, nmig,
                            thetai, mi, float(qmax), float(mprior), int(expo), int(first),
                            int(last), _dp(pk))
    return float(pk[0]), v


def greater_than_one(row, npsize, nmig, kind, i, j, prior_max):
    """One genealogy's term of gtpops (kind 0) or gtmig (kind 1), gtint.cpp:127-331."""
    row = np.ascontiguousarray(row, np.float32)
    f = lib.ora_gtmig_one if kind else lib.ora_gtpops_one
    return f(row.ctypes.data_as(C.POINTER(C.c_float)), npsize, nmig, i, j, float(prior_max))


def greater_than(table, npsize, nmig, kind, i, j, prior_max):
    """P(param i > param j) over the table with print_greater_than_tests's
    subsampling (gtint.cpp:335-349)."""
    table = np.ascontiguousarray(table, np.float32)
    G, K = table.shape
    return lib.ora_greater_than(table.ctypes.data_as(C.POINTER(C.c_float)), G, K, npsize, nmig,
                                int(kind), i, j, float(prior_max))


lib.ora_swapchains.restype = C.c_int
lib.ora_swapchains.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                               C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int),
                               C.POINTER(C.c_int)]


def swapchains(S, beta, swaptries, uniforms):
    S = np.ascontiguousarray(S, np.float64)
    b = np.array(beta, np.float64)
    uni = np.ascontiguousarray(uniforms, np.float64)
    used = np.zeros(1, np.int32)
    pairs = np.zeros(3 * max(swaptries, 1), np.int32)
    acc = lib.ora_swapchains(len(S), _dp(S), _dp(b), swaptries, _dp(uni), len(uni),
                             _ip(used), _ip(pairs))
    return b, acc, pairs[: 3 * swaptries].reshape(-1, 3), int(used[0])


for _f in ("ora_uppergamma", "ora_lowergamma"):
    getattr(lib, _f).restype = C.c_double
    getattr(lib, _f).argtypes = [C.c_int, C.c_double]
lib.ora_logfact.restype = C.c_double
lib.ora_logfact.argtypes = [C.c_int]
lib.ora_integrate_coalescent_term.restype = C.c_double
lib.ora_integrate_coalescent_term.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                              C.c_double]
lib.ora_integrate_migration_term.restype = C.c_double
lib.ora_integrate_migration_term.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
lib.ora_treeweight.restype = C.c_int
lib.ora_treeweight.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                               C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int, C.c_int,
                               C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                               C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                               C.POINTER(C.c_int), C.c_double, C.POINTER(C.c_int),
                               C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int),
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]


def uppergamma(a, x):
    return lib.ora_uppergamma(int(a), float(x))


def lowergamma(a, x):
    return lib.ora_lowergamma(int(a), float(x))


def integrate_coalescent_term(cc, fc, hcc, qmax, qmin=0.0):
    return lib.ora_integrate_coalescent_term(int(cc), float(fc), float(hcc), float(qmax),
                                             float(qmin))


def integrate_migration_term(cm, fm, mmax, mmin=0.0):
    return lib.ora_integrate_migration_term(int(cm), float(fm), float(mmax), float(mmin))


def treeweight(g, model, tvals, samppop, hval=1.0):
    """Genealogy weights of one genealogy: dict of cc, fc, hcc [nper, P] and
    mc, fm [nper, P, P], plus the total length."""
    from ima2p_amd.immodel import mig_arrays
    t = Tree(g)
    P, nper = model.npops, model.nper
    me, mt, mto = mig_arrays(g)
    pop = np.ascontiguousarray(g.pop, np.int32)
    plist = np.ascontiguousarray(model.plist, np.int32)
    npres = np.ascontiguousarray(model.npresent, np.int32)
    pdown = np.ascontiguousarray(model.popdown, np.int32)
    pope = np.ascontiguousarray(model.pope, np.int32)
    join = np.ascontiguousarray(model.join, np.int32)
    tv = np.ascontiguousarray(np.append(tvals, 0.0), np.float64)
    sp = np.ascontiguousarray(samppop, np.int32)
    cc = np.zeros(nper * P, np.int32)
    fc = np.zeros(nper * P)
    hcc = np.zeros(nper * P)
    mc = np.zeros(nper * P * P, np.int32)
    fm = np.zeros(nper * P * P)
    length = np.zeros(1)
    rc = lib.ora_treeweight(t.ref, _ip(pop), len(me), _ip(me), _dp(mt), _ip(mto), P, nper,
                            _ip(plist), _ip(npres), _ip(pdown), _ip(pope), _ip(join), _dp(tv),
                            _ip(sp), float(hval), _ip(cc), _dp(fc), _dp(hcc), _ip(mc), _dp(fm),
                            _dp(length))
    if rc == -2:
        raise ValueError("treeweight: a population label does not fit the model's periods")
    if rc != 0:
        raise ValueError("treeweight: event count mismatch")
    return {"cc": cc.reshape(nper, P), "fc": fc.reshape(nper, P), "hcc": hcc.reshape(nper, P),
            "mc": mc.reshape(nper, P, P), "fm": fm.reshape(nper, P, P), "length": float(length[0])}


def integrate_tree_prob(weights_list, model):
    """integrate_tree_prob over the sum of several genealogies' weights
    (allgweight): returns (probg, qintegrate, mintegrate)."""
    cc = sum(w["cc"] for w in weights_list)
    fc = sum(w["fc"] for w in weights_list)
    hcc = sum(w["hcc"] for w in weights_list)
    mc = sum(w["mc"] for w in weights_list)
    fm = sum(w["fm"] for w in weights_list)
    q, m = [], []
    for i, wp in enumerate(model.theta_wp):
        c = sum(int(cc[k, r]) for k, r in wp)
        f = 0.0
        h = 0.0
        for k, r in wp:
            f += fc[k, r]
            h += hcc[k, r]
        q.append(integrate_coalescent_term(c, f, h, model.qmax[i]))
    for i, wp in enumerate(model.mig_wp):
        c = sum(int(mc[k, a, b]) for k, a, b in wp)
        f = 0.0
        for k, a, b in wp:
            f += fm[k, a, b]
        m.append(integrate_migration_term(c, f, model.mmax[i]))
    return float(sum(q) + sum(m)), q, m


def gsampinf_row(wall, integrals, pdg, probg, tvals, model):
    """savegsampinf (ginfo.cpp:319-377) for one chain: wall is the chain's
    all-loci weight record (cc[P][P] | fc[P][P] | hcc[P][P] | mc[P][P][P] |
    fm[P][P][P] | length, as doubles), integrals its qintegrate then
    mintegrate, pdg its loci' log-likelihoods (summed in locus order),
    probg, tvals the split times.  Per population-size parameter: the int
    sum of cc and the float sums of fc and hcc over its wp positions in
    order; per migration parameter: the int sum of mc and the float sum of
    fm; then qintegrate, mintegrate, pdg, probg, t -- all float32, in the
    layout of ginfo.cpp:288-316."""
    P = model.npops
    P2, P3 = P * P, P * P * P
    w = np.asarray(wall, np.float64)
    cc, fc, hcc = w[:P2].reshape(P, P), w[P2:2 * P2].reshape(P, P), w[2 * P2:3 * P2].reshape(P, P)
    mc = w[3 * P2:3 * P2 + P3].reshape(P, P, P)
    fm = w[3 * P2 + P3:3 * P2 + 2 * P3].reshape(P, P, P)
    f32 = np.float32
    ccs, fcs, hcs, mcs, fms = [], [], [], [], []
    for wp in model.theta_wp:
        c, f, h = 0, f32(0.0), f32(0.0)
        for k, r in wp:
            c += int(cc[k, r])
            f = f32(f + f32(fc[k, r]))
            h = f32(h + f32(hcc[k, r]))
        ccs.append(f32(c))
        fcs.append(f)
        hcs.append(h)
    for wp in model.mig_wp:
        c, f = 0, f32(0.0)
        for k, a, b in wp:
            c += int(mc[k, a, b])
            f = f32(f + f32(fm[k, a, b]))
        mcs.append(f32(c))
        fms.append(f)
    nps = len(model.theta_wp)
    integ = [f32(x) for x in integrals]
    pd = 0.0
    for x in pdg:
        pd += float(x)
    row = ccs + fcs + hcs + mcs + fms + integ[:nps] + integ[nps:] + [f32(pd), f32(probg)] + \
        [f32(t) for t in list(tvals)[:P - 1]]
    return np.array(row, np.float32)


def updategenealogy_decisions(pdg, weights, new_pdg, new_weights, ok, hastings, uniforms,
                              beta, model, gbeta=1.0):
    """The accept step of updategenealogy (update_gtree.cpp:900-961) for one
    chain's loci taken in order, as qupdate calls it (ima_main_mpi.cpp:
    1821-1841): the chain's all-loci weights are the sum of its loci'
    weights (allgweight); locus l's proposal has log-likelihood new_pdg[l],
    weights new_weights[l] and Hastings term hastings[l] (migweight +
    slideweight + Atermsum); ok[l] is False for a rejected infinite-sites or
    zero stepwise likelihood (rejectIS).  tpw = probg' - probg with probg'
    from integrate_tree_prob over the weights with locus l's replaced;
    accept iff U < min(1, exp(beta (tpw + gbeta (new_pdg - pdg)) + h)).
    Returns (decisions, pdg, weights, probg) after the last locus."""
    pdg = list(pdg)
    weights = list(weights)
    probg = integrate_tree_prob(weights, model)[0]
    dec = []
    for l in range(len(pdg)):
        acc = False
        if ok[l]:
            trial = weights[:l] + [new_weights[l]] + weights[l + 1:]
            pn = integrate_tree_prob(trial, model)[0]
            x = beta * ((pn - probg) + gbeta * (new_pdg[l] - pdg[l])) + hastings[l]
            acc = uniforms[l] < min(1.0, float(np.exp(x)))
            if acc:
                weights = trial
                pdg[l] = new_pdg[l]
                probg = pn
        dec.append(bool(acc))
    return dec, pdg, weights, probg


def getmprob_nw(period, lastp, mrate, mtime, mcount, uppop, dpop, cm2pop, numpops):
    return lib.ora_getmprob_nw(int(period), int(lastp), float(mrate), float(mtime), int(mcount),
                               int(uppop), int(dpop), int(cm2pop), int(numpops))


def poisson_cond(param, cond, key, ctr=0):
    return lib.ora_poisson_cond(float(param), int(cond), int(key) & (2**64 - 1), int(ctr))


def update_mig_tnw(g, model, tv, period, newt, key, ctr=0, cap=None):
    """changet_NW's migration update of Genealogy g (update_mig_tNW,
    update_t_NW.cpp:392-792) for split `period` moving from tv[period] to
    newt, edges drawing from Philox streams keyed by (key, edge) at counter
    ctr.  Returns (rc, updated, migweight, new Genealogy); the input is left
    as it is.  cap bounds the new migration count (default: ample)."""
    from ima2p_amd.immodel import mig_arrays
    me, mt, mo = mig_arrays(g)
    nm = len(me)
    if cap is None:
        cap = 4 * nm + 4096
    ne = 2 * g.n - 1
    me_ = np.zeros(cap + 1, np.int32)
    mt_ = np.zeros(cap + 1, np.float64)
    mo_ = np.zeros(cap + 1, np.int32)
    me_[:nm], mt_[:nm], mo_[:nm] = me, mt, mo
    pop = np.ascontiguousarray(g.pop, np.int32).copy()
    nmig = np.array([nm], np.int32)
    arrs = [np.ascontiguousarray(a, np.int32) for a in
            (model.plist, model.npresent, model.popdown, model.pope, model.join)]
    w = np.zeros(1)
    upd = np.zeros(1, np.int32)
    up0, up1, dn = (np.ascontiguousarray(a, np.int32) for a in (g.up0, g.up1, g.down))
    tm = np.ascontiguousarray(g.time, np.float64)
    tvv = np.ascontiguousarray(tv, np.float64)
    rc = lib.ora_update_mig_tnw(g.n, _ip(up0), _ip(up1), _ip(dn), _dp(tm), int(g.root), _ip(pop),
                                _ip(nmig), int(cap), _ip(me_), _dp(mt_), _ip(mo_), model.npops,
                                *[_ip(a) for a in arrs], _dp(tvv), int(period), float(newt),
                                int(key) & (2**64 - 1), int(ctr), _dp(w), _ip(upd))
    out = g.copy()
    if rc == 0 and upd[0]:
        out.pop = pop.copy()
        out.mig = [[] for _ in range(ne)]
        for j in range(int(nmig[0])):
            out.mig[int(me_[j])].append((float(mt_[j]), int(mo_[j])))
    return int(rc), int(upd[0]), float(w[0]), out


def mc3_bench(loci, u, kappa, samppop, model, chains, tvs, beta, tmax, seed=1,
              seconds=1.0, max_sweeps=0, scale_width=0.2, ry1=False, urate=False,
              rates=None, topo=False, nw=0, mig_room=256):
    """The CPU baseline's C sweep over chains[c][l] (Genealogy objects with
    population labels and migrations; updated in place) with split times
    tvs[c] (a float array [nchains, P-1], updated in place).  ry1: the
    alternate-sweep time update is changet_RY1's rubber band instead of the
    global rescale.  urate: changeu after every fifth sweep.  u (per locus
    or [nchains, nloci]) and kappa (scalar or [nchains, nloci]) start each
    chain; rates=(u, kappa), two float64 [nchains, nloci] arrays, take their
    place and receive the final values.  topo: the genealogy update is the
    reference's own move (slider_nomigration without migration parameters,
    slider and addmigration with them; topologies, labels, roots and
    migrations of the Genealogy objects are updated).  nw (models
    with migration): 1, each time update is changet_NW with probability 1/2
    (the reference's choice between RY1 and NW); 2, every sweep is one NW
    update per chain and nothing else; populations and migrations of the
    Genealogy objects are updated, each genealogy holding at most its
    migration count plus max(mig_room, that count).  Returns
    (sweeps, stats[6] =
    genealogy proposals, accepts, time-update proposals, accepts,
    mutation-rate steps, accepts)."""
    from ima2p_amd.immodel import mig_arrays
    L = len(loci)
    P = model.npops
    n_l = np.array([chains[0][l].n for l in range(L)], np.int32)
    S = np.array([lc.seq.shape[1] for lc in loci], np.int32)
    seq = np.concatenate([np.ascontiguousarray(lc.seq, np.int32).ravel() for lc in loci])
    seq_off = np.cumsum([0] + [lc.seq.size for lc in loci])[:-1].astype(np.int32)
    mult = np.concatenate([np.ascontiguousarray(lc.mult, np.int32) for lc in loci])
    mult_off = np.cumsum([0] + [len(lc.mult) for lc in loci])[:-1].astype(np.int32)
    tot = np.array([lc.totsites for lc in loci], np.int32)
    pi = np.concatenate([np.asarray(lc.pi, float) for lc in loci])
    nch = len(chains)
    if rates is None:
        uu = np.ascontiguousarray(np.broadcast_to(np.asarray(u, float), (nch, L)), float)
        kk = np.ascontiguousarray(np.broadcast_to(np.asarray(kappa, float), (nch, L)), float)
    else:
        uu, kk = rates
        assert all(a.dtype == np.float64 and a.flags.c_contiguous and a.shape == (nch, L)
                   for a in (uu, kk))
    sp = np.ascontiguousarray(np.broadcast_to(np.asarray(samppop, np.int32), (L, P)), np.int32)
    toff = np.cumsum([0] + [len(w) for w in model.theta_wp]).astype(np.int32)
    twp = np.array([x for w in model.theta_wp for kv in w for x in kv] or [0], np.int32)
    moff = np.cumsum([0] + [len(w) for w in model.mig_wp]).astype(np.int32)
    mwp = np.array([x for w in model.mig_wp for kv in w for x in kv] or [0], np.int32)
    gens = [g for row in chains for g in row]
    up0 = np.concatenate([g.up0 for g in gens]).astype(np.int32)
    up1 = np.concatenate([g.up1 for g in gens]).astype(np.int32)
    dn = np.concatenate([g.down for g in gens]).astype(np.int32)
    pop = np.concatenate([g.pop for g in gens]).astype(np.int32)
    tm = np.concatenate([g.time for g in gens]).astype(np.float64)
    roots = np.array([g.root for g in gens], np.int32)
    migs = [mig_arrays(g) for g in gens]
    nm = np.array([len(m[0]) for m in migs], np.int32)
    roomy = bool(nw) or (bool(topo) and model.nmig > 0)
    if roomy:
        caps = np.array([k + max(int(mig_room), k) for k in nm], np.int32)
        room = [np.zeros(int(c - k), np.int32) for c, k in zip(caps, nm)]
        me = np.concatenate([x for m, r in zip(migs, room) for x in (m[0], r)]
                            + [np.zeros(1, np.int32)]).astype(np.int32)
        mt = np.concatenate([x for m, r in zip(migs, room) for x in (m[1], r.astype(float))]
                            + [np.zeros(1)]).astype(np.float64)
        mo = np.concatenate([x for m, r in zip(migs, room) for x in (m[2], r)]
                            + [np.zeros(1, np.int32)]).astype(np.int32)
    else:
        caps = None
        me = np.concatenate([m[0] for m in migs] + [np.zeros(1, np.int32)]).astype(np.int32)
        mt = np.concatenate([m[1] for m in migs] + [np.zeros(1)]).astype(np.float64)
        mo = np.concatenate([m[2] for m in migs] + [np.zeros(1, np.int32)]).astype(np.int32)
    tv = tvs if P > 1 else np.zeros(1)
    assert P == 1 or (tv.dtype == np.float64 and tv.flags.c_contiguous and tv.size == len(chains) * (P - 1))
    b = np.ascontiguousarray(beta, float)
    qm = np.ascontiguousarray(model.qmax if model.npsize else [1.0], float)
    mm = np.ascontiguousarray(model.mmax if model.nmig else [1.0], float)
    arrs = [np.ascontiguousarray(a, np.int32) for a in
            (model.plist, model.npresent, model.popdown, model.pope, model.join)]
    stats = (C.c_long * 6)()
    roots_out = np.zeros(len(gens), np.int32)
    sw = lib.ora_mc3_bench(len(chains), L, _ip(n_l), _ip(S), _ip(seq), _ip(seq_off), _ip(mult),
                           _ip(mult_off), _ip(tot), _dp(pi), _dp(uu), _dp(kk), _ip(sp), P,
                           *[_ip(a) for a in arrs], model.npsize, _ip(toff), _ip(twp),
                           model.nmig, _ip(moff), _ip(mwp), _dp(qm), _dp(mm), float(tmax),
                           _ip(up0), _ip(up1), _ip(dn), _ip(pop), _dp(tm), _ip(roots), _ip(nm),
                           _ip(me), _dp(mt), _ip(mo), _dp(tv), _dp(b), float(scale_width),
                           int(seed), float(seconds), int(max_sweeps), stats, int(bool(ry1)),
                           int(bool(urate)), int(bool(topo)), _ip(roots_out), int(nw),
                           _ip(caps) if caps is not None else None)
    if sw < 0:
        raise ValueError("mc3_bench: bad arguments")
    off = moff_ = 0
    for gi, (g, m) in enumerate(zip(gens, migs)):
        ne = 2 * g.n - 1
        g.time[:] = tm[off:off + ne]
        if topo:
            g.up0[:] = up0[off:off + ne]
            g.up1[:] = up1[off:off + ne]
            g.down[:] = dn[off:off + ne]
            g.pop[:] = pop[off:off + ne]
            g.root = int(roots_out[gi])
        if roomy:
            g.pop[:] = pop[off:off + ne]
            g.mig = [[] for _ in range(ne)]
            for j in range(int(nm[gi])):
                q = moff_ + j
                g.mig[int(me[q])].append((float(mt[q]), int(mo[q])))
            off += ne
            moff_ += int(caps[gi])
            continue
        off += ne
        k = len(m[0])
        times = mt[moff_:moff_ + k]
        j = 0
        for edge in range(ne):
            lst = sorted(g.mig[edge])
            g.mig[edge] = [(float(times[j + i]), q) for i, (_, q) in enumerate(lst)]
            j += len(lst)
        moff_ += k
    return int(sw), list(stats)
