"""Host-side front end of the HIP engine (thin layer over the C ABI).

``Engine`` owns one context (one HIP stream) on one device.  ``RefShim``
mirrors the reference's per-(chain, locus) entry points
(``likelihoodHKY``/``likelihoodIS``/``likelihoodSW``, imamp.hpp:1167-1174,
and ``swapchains``, swapchains.cpp:527) over genealogy slots, so callers
written against the reference read the same here.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class EngineError(RuntimeError):
    pass


def _i32(a):
    a = np.ascontiguousarray(a, np.int32)
    return a, a.ctypes.data_as(L.i32p)


def _f64(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(L.dp)


class Engine:
    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.ima_ctx_create(device, C.byref(h))
        if rc != 0:
            raise EngineError(f"ima_ctx_create failed with code {rc}")
        self.h = h
        self.loci = []          # (model, n, S)

    def close(self):
        if self.h:
            self.lib.ima_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.lib.ima_ctx_last_error(self.h).decode(errors="replace")
            err = EngineError(f"{what}: code {rc}: {msg}")
            err.code = rc
            raise err

    # -------------------------------------------------------------- loci --
    def add_locus_hky(self, loc) -> int:
        seq = np.ascontiguousarray(loc.seq, np.int8)
        mult, mp = _i32(loc.mult)
        pi, pp = _f64(loc.pi)
        out = C.c_int()
        self._chk(self.lib.ima_locus_add_hky(self.h, seq.shape[0], seq.shape[1],
                                             seq.ctypes.data_as(L.i8p), mp, int(loc.totsites),
                                             pp, C.byref(out)), "add_locus_hky")
        self.loci.append((L.MODEL_HKY, seq.shape[0], seq.shape[1]))
        return out.value

    def add_locus_is(self, loc) -> int:
        seq = np.ascontiguousarray(loc.seq, np.int8)
        out = C.c_int()
        self._chk(self.lib.ima_locus_add_is(self.h, seq.shape[0], seq.shape[1],
                                            seq.ctypes.data_as(L.i8p), C.byref(out)),
                  "add_locus_is")
        self.loci.append((L.MODEL_IS, seq.shape[0], seq.shape[1]))
        return out.value

    def add_locus_sw(self, n: int) -> int:
        out = C.c_int()
        self._chk(self.lib.ima_locus_add_sw(self.h, int(n), C.byref(out)), "add_locus_sw")
        self.loci.append((L.MODEL_SW, int(n), 0))
        return out.value

    def set_sumlogk(self, locus, v):
        self._chk(self.lib.ima_locus_set_sumlogk(self.h, locus, float(v)), "set_sumlogk")

    def sumlogk_from_slot(self, locus, slot) -> float:
        out = C.c_double()
        self._chk(self.lib.ima_locus_sumlogk_from_slot(self.h, locus, slot, C.byref(out)),
                  "sumlogk_from_slot")
        return out.value

    # ------------------------------------------------------------- slots --
    def alloc_slots(self, loci) -> int:
        la, lp = _i32(loci)
        first = C.c_int()
        self._chk(self.lib.ima_slots_alloc(self.h, len(la), lp, C.byref(first)), "alloc_slots")
        return first.value

    def slot_info(self, slot):
        locus, n = C.c_int(), C.c_int()
        self._chk(self.lib.ima_slot_info(self.h, slot, C.byref(locus), C.byref(n)), "slot_info")
        return locus.value, n.value

    def write_slots(self, first, gens, A=None):
        up0, p0 = _i32(np.concatenate([g.up0 for g in gens]))
        up1, p1 = _i32(np.concatenate([g.up1 for g in gens]))
        dn, pd = _i32(np.concatenate([g.down for g in gens]))
        tm, pt = _f64(np.concatenate([g.time for g in gens]))
        rt, pr = _i32([g.root for g in gens])
        if A is not None:
            aa, pa = _i32(np.concatenate([np.asarray(a) for a in A]))
        else:
            aa, pa = None, None
        self._chk(self.lib.ima_slots_write(self.h, first, len(gens), p0, p1, pd, pt, pr, pa),
                  "write_slots")

    def read_slots(self, first, count):
        n = [self.slot_info(first + i)[1] for i in range(count)]
        tot = sum(2 * k - 1 for k in n)
        up0 = np.zeros(tot, np.int32)
        up1 = np.zeros(tot, np.int32)
        dn = np.zeros(tot, np.int32)
        tm = np.zeros(tot)
        A = np.zeros(tot, np.int32)
        rt = np.zeros(count, np.int32)
        self._chk(self.lib.ima_slots_read(self.h, first, count, up0.ctypes.data_as(L.i32p),
                                          up1.ctypes.data_as(L.i32p), dn.ctypes.data_as(L.i32p),
                                          tm.ctypes.data_as(L.dp), rt.ctypes.data_as(L.i32p),
                                          A.ctypes.data_as(L.i32p)), "read_slots")
        return up0, up1, dn, tm, rt, A

    def read_pops(self, first, count):
        n = [self.slot_info(first + i)[1] for i in range(count)]
        pop = np.zeros(sum(2 * k - 1 for k in n), np.int32)
        self._chk(self.lib.ima_slots_read_pops(self.h, first, count, pop.ctypes.data_as(L.i32p)),
                  "read_pops")
        return pop

    # ---------------------------------------------------- genealogy prior --
    def model_add(self, model):
        """Register an immodel.IMModel; returns its id."""
        P = model.npops
        toff = np.cumsum([0] + [len(w) for w in model.theta_wp]).astype(np.int32)
        twp = np.array([x for w in model.theta_wp for kv in w for x in kv] or [0], np.int32)
        moff = np.cumsum([0] + [len(w) for w in model.mig_wp]).astype(np.int32)
        mwp = np.array([x for w in model.mig_wp for kv in w for x in kv] or [0], np.int32)
        arrs = [np.ascontiguousarray(a, np.int32) for a in
                (model.plist, model.npresent, model.popdown, model.pope, model.join)]
        qm, pq = _f64(model.qmax if model.npsize else [1.0])
        mm, pm = _f64(model.mmax if model.nmig else [1.0])
        out = C.c_int(-1)
        self._chk(self.lib.ima_model_add(
            self.h, P, *[a.ctypes.data_as(L.i32p) for a in arrs], model.npsize,
            toff.ctypes.data_as(L.i32p), twp.ctypes.data_as(L.i32p), model.nmig,
            moff.ctypes.data_as(L.i32p), mwp.ctypes.data_as(L.i32p), pq, pm, C.byref(out)),
            "model_add")
        return out.value

    def locus_set_model(self, locus, model_id, samppop, hval=1.0):
        sp, psp = _i32(samppop)
        self._chk(self.lib.ima_locus_set_model(self.h, locus, model_id, psp, float(hval)),
                  "locus_set_model")

    def write_pops(self, first, gens):
        """Population labels and migration events (Genealogy.pop / .mig)."""
        from .immodel import mig_arrays
        pop, pp = _i32(np.concatenate([g.pop for g in gens]))
        es, ts, tos, cnt = [], [], [], []
        for g in gens:
            e, t, to = mig_arrays(g)
            es.append(e)
            ts.append(t)
            tos.append(to)
            cnt.append(len(e))
        nm, pn = _i32(cnt)
        me, pme = _i32(np.concatenate(es) if sum(cnt) else [0])
        mt, pmt = _f64(np.concatenate(ts) if sum(cnt) else [0.0])
        mo, pmo = _i32(np.concatenate(tos) if sum(cnt) else [0])
        self._chk(self.lib.ima_slots_write_pops(self.h, first, len(gens), pp, pn, pme, pmt, pmo),
                  "write_pops")

    def slots_reserve_migs(self, first, count, cap):
        self._chk(self.lib.ima_slots_reserve_migs(self.h, first, count, int(cap)),
                  "slots_reserve_migs")

    def t_nw_update(self, slots, tvals, period, newt, seed, counter=0):
        """changet_NW's migration update of a chain's genealogies (slots) for
        split `period` moving from tvals[period] to newt; returns
        (migweight[n], updated[n])."""
        sl, psl = _i32(slots)
        tv, ptv = _f64(tvals)
        w = np.zeros(len(sl))
        upd = np.zeros(len(sl), np.int32)
        self._chk(self.lib.ima_t_nw_update(self.h, psl, len(sl), ptv, int(period), float(newt),
                                           int(seed) & (2**64 - 1), int(counter),
                                           w.ctypes.data_as(L.dp), upd.ctypes.data_as(L.i32p)),
                  "t_nw_update")
        return w, upd

    def treeweight(self, slots, tvals, P):
        """Per-slot weights: dict of cc/fc/hcc [n, P, P], mc/fm [n, P, P, P], length."""
        slots = np.ascontiguousarray(slots, np.int32)
        k = len(slots)
        if P > 1:
            tv = np.asarray(tvals, float).reshape(-1, P - 1)
            tv = np.ascontiguousarray(np.broadcast_to(tv, (k, P - 1)) if len(tv) == 1 else tv)
        else:
            tv = np.zeros((k, 1))
        cc = np.zeros((k, P, P), np.int32)
        fc = np.zeros((k, P, P))
        hcc = np.zeros((k, P, P))
        mc = np.zeros((k, P, P, P), np.int32)
        fm = np.zeros((k, P, P, P))
        ln = np.zeros(k)
        self._chk(self.lib.ima_treeweight(
            self.h, slots.ctypes.data_as(L.i32p), k, tv.ctypes.data_as(L.dp),
            cc.ctypes.data_as(L.i32p), fc.ctypes.data_as(L.dp), hcc.ctypes.data_as(L.dp),
            mc.ctypes.data_as(L.i32p), fm.ctypes.data_as(L.dp), ln.ctypes.data_as(L.dp)),
            "treeweight")
        return {"cc": cc, "fc": fc, "hcc": hcc, "mc": mc, "fm": fm, "length": ln}

    def prior(self, groups, tvals, model):
        """integrate_tree_prob over each group of slots; tvals[g] its split
        times.  Returns (probg[G], qint[G, npsize], mint[G, nmig])."""
        slots = np.ascontiguousarray(np.concatenate([np.asarray(g, np.int32) for g in groups]),
                                     np.int32)
        off = np.cumsum([0] + [len(g) for g in groups]).astype(np.int32)
        G, P = len(groups), model.npops
        tv = np.ascontiguousarray(np.asarray(tvals, float).reshape(G, max(P - 1, 1))
                                  if P > 1 else np.zeros((G, 1)))
        probg = np.zeros(G)
        q = np.zeros((G, max(model.npsize, 1)))
        m = np.zeros((G, max(model.nmig, 1)))
        self._chk(self.lib.ima_prior(self.h, slots.ctypes.data_as(L.i32p),
                                     off.ctypes.data_as(L.i32p), G, tv.ctypes.data_as(L.dp),
                                     probg.ctypes.data_as(L.dp), q.ctypes.data_as(L.dp),
                                     m.ctypes.data_as(L.dp)), "prior")
        return probg, q[:, :model.npsize], m[:, :model.nmig]

    # ----------------------------------------------------------- scoring --
    @staticmethod
    def items(slots, u, kappa=None, tr=None):
        slots = np.asarray(slots)
        arr = (L.Item * len(slots))()
        u = np.broadcast_to(np.asarray(u, float), slots.shape)
        kappa = np.broadcast_to(np.asarray(1.0 if kappa is None else kappa, float), slots.shape)
        tr = np.broadcast_to(np.asarray(1.0 if tr is None else tr, float), slots.shape)
        for i in range(len(slots)):
            arr[i].slot = int(slots[i])
            arr[i].flags = 0
            arr[i].u = float(u[i])
            arr[i].kappa = float(kappa[i])
            arr[i].tr = float(tr[i])
        return arr

    def score(self, items, precision=L.PREC_FP64):
        n = len(items)
        out = np.zeros(n)
        st = np.zeros(n, np.int32)
        self._chk(self.lib.ima_score(self.h, items, n, precision, out.ctypes.data_as(L.dp),
                                     st.ctypes.data_as(L.i32p)), "score")
        return out, st

    def score_repeat(self, items, reps, precision=L.PREC_FP64):
        """The same scoring `reps` times back to back (kernel timing)."""
        n = len(items)
        out = np.zeros(n)
        st = np.zeros(n, np.int32)
        self._chk(self.lib.ima_score_repeat(self.h, items, n, precision, int(reps),
                                            out.ctypes.data_as(L.dp),
                                            st.ctypes.data_as(L.i32p)), "score_repeat")
        return out, st

    def is_site_nodes(self, slot, S):
        out = np.zeros(max(S, 1), np.int32)
        self._chk(self.lib.ima_is_site_nodes(self.h, slot, out.ctypes.data_as(L.i32p)),
                  "is_site_nodes")
        return out[:S]

    def sw_dlike(self, slot, u, tr=1.0):
        it = self.items([slot], u, tr=tr)
        n = self.slot_info(slot)[1]
        out = np.zeros(2 * n - 1)
        self._chk(self.lib.ima_sw_dlike(self.h, it, out.ctypes.data_as(L.dp)), "sw_dlike")
        return out

    def sw_update_A(self, moves, calc=True):
        """finishSWupdateA for a batch of moves (objects with the fields of
        ima_sw_move, e.g. sim.SWMove with .slot set).  Returns (dlike,
        Aterm, newA, edge_dlikeA[n, 4])."""
        n = len(moves)
        arr = (L.SWMove * max(n, 1))()
        for i, m in enumerate(moves):
            arr[i] = L.SWMove(m.slot, m.edge, m.downedge, m.sisedge, m.newsisedge, m.copyA0,
                              m.copyA1, int(m.down_was_root), m.hold_down_A, m.minA, m.maxA,
                              int(m.flip), (C.c_double * 3)(*m.copy_dlikeA), m.newsis_dlikeA,
                              m.u, m.uni)
        dl, at, ed = np.zeros(n), np.zeros(n), np.zeros((n, 4))
        na = np.zeros(n, np.int32)
        self._chk(self.lib.ima_sw_update_A(self.h, C.cast(arr, C.c_void_p), n, int(bool(calc)),
                                           dl.ctypes.data_as(L.dp), at.ctypes.data_as(L.dp),
                                           na.ctypes.data_as(L.i32p), ed.ctypes.data_as(L.dp)),
                  "sw_update_A")
        return dl, at, na, ed

    def likelihood_hky(self, locus, g, u, kappa, e=(-1, -1, -1, -1)):
        up0, p0 = _i32(g.up0)
        up1, p1 = _i32(g.up1)
        dn, pd = _i32(g.down)
        tm, pt = _f64(g.time)
        out = C.c_double()
        self._chk(self.lib.ima_likelihood_hky(self.h, locus, p0, p1, pd, pt, int(g.root),
                                              float(u), float(kappa), *[int(x) for x in e],
                                              C.byref(out)), "likelihood_hky")
        return out.value

    # -------------------------------------------------------------- swap --
    def swap(self, S, beta, swaptries, seed, counter):
        S, ps = _f64(S)
        b = np.array(beta, np.float64)
        pairs = np.zeros(3 * max(swaptries, 1), np.int32)
        nacc = C.c_int()
        self._chk(self.lib.ima_swap(self.h, len(S), ps, b.ctypes.data_as(L.dp), swaptries,
                                    seed, counter, C.byref(nacc), pairs.ctypes.data_as(L.i32p)),
                  "swap")
        return b, nacc.value, pairs[: 3 * swaptries].reshape(-1, 3)

    # ------------------------------------------------------------ L mode --
    def lmode_table(self, table, npsize, nmig, npops) -> int:
        t = np.ascontiguousarray(table, np.float32)
        out = C.c_int()
        self._chk(self.lib.ima_lmode_table(self.h, t.ctypes.data_as(L.fp), t.shape[0],
                                           t.shape[1], npsize, nmig, npops, C.byref(out)),
                  "lmode_table")
        return out.value

    def lmode_table_free(self, tid):
        self._chk(self.lib.ima_lmode_table_free(self.h, int(tid)), "lmode_table_free")

    def lmode_table_reserve(self, capacity, npsize, nmig, npops) -> int:
        """An empty resident table the device sweep appends rows to."""
        out = C.c_int()
        self._chk(self.lib.ima_lmode_table_reserve(self.h, int(capacity), int(npsize), int(nmig),
                                                   int(npops), C.byref(out)), "lmode_table_reserve")
        return out.value

    def lmode_table_rows(self, tid):
        G, K = C.c_int(), C.c_int()
        self._chk(self.lib.ima_lmode_table_info(self.h, tid, C.byref(G), C.byref(K), None, None,
                                                None), "lmode_table_info")
        return G.value, K.value

    def lmode_table_read(self, tid, first=0, count=None):
        G, K = self.lmode_table_rows(tid)
        if count is None:
            count = G - first
        out = np.zeros((count, K), np.float32)
        self._chk(self.lib.ima_lmode_table_read(self.h, tid, int(first), int(count),
                                                out.ctypes.data_as(L.fp)), "lmode_table_read")
        return out

    def mc3_record(self, tid):
        """Append the cold chain's genealogy summary row (savegsampinf)."""
        self._chk(self.lib.ima_mc3_record(self.h, int(tid)), "mc3_record")

    def mc3_read_weights(self, R, nparam):
        """Each chain's all-loci weight record [C, R] and integrals [C, nparam]."""
        C_, _ = self.mc3_shape
        w = np.zeros((C_, R))
        q = np.zeros((C_, max(nparam, 1)))
        self._chk(self.lib.ima_mc3_read_weights(self.h, w.ctypes.data_as(L.dp),
                                                q.ctypes.data_as(L.dp)), "mc3_read_weights")
        return w, q[:, :nparam]

    def lmode_free(self, tid):
        self._chk(self.lib.ima_lmode_table_free(self.h, tid), "lmode_free")

    def jointp(self, tid, x, mode=L.LMODE_EXACT, modeltype=0, prange=None):
        x = np.ascontiguousarray(np.atleast_2d(x), np.float64)
        npnt, np_ = x.shape
        r0, r1 = prange if prange is not None else (0, np_)
        out = np.zeros(npnt)
        ess = np.zeros(npnt)
        self._chk(self.lib.ima_lmode_jointp(self.h, tid, x.ctypes.data_as(L.dp), npnt,
                                            modeltype, r0, r1, mode,
                                            out.ctypes.data_as(L.dp), ess.ctypes.data_as(L.dp)),
                  "jointp")
        return out, ess

    def joint_peak(self, tid, lower, upper, modeltype=0, xmap=None, mult=100, seed=1,
                   mode=L.LMODE_REF_COMPAT):
        """Differential-evolution joint peak (ima_lmode_joint_peak).
        Returns (peak vector, log P at the peak, ESS, jointp evaluations)."""
        lo, plo = _f64(lower)
        hi, phi = _f64(upper)
        peak = np.zeros(len(lo))
        lp, es = C.c_double(), C.c_double()
        ev = C.c_int64()
        xm = None if xmap is None else np.ascontiguousarray(xmap, np.float64)
        self._chk(self.lib.ima_lmode_joint_peak(
            self.h, int(tid), int(modeltype), None if xm is None else xm.ctypes.data_as(L.dp),
            plo, phi, int(mult), int(seed), int(mode), peak.ctypes.data_as(L.dp), C.byref(lp),
            C.byref(es), C.byref(ev)), "joint_peak")
        return peak, lp.value, es.value, ev.value

    def marginal_peak(self, tid, param, lower, upper, expo=False, mean=1.0):
        """(peak, log density at the peak, lower 95%, upper 95%, no-migration
        2LLR) of one parameter's marginal posterior (ima_lmode_marginal_peak)."""
        out = np.zeros(5)
        self._chk(self.lib.ima_lmode_marginal_peak(self.h, int(tid), int(param), float(lower),
                                                   float(upper), int(expo), float(mean),
                                                   out.ctypes.data_as(L.dp)), "marginal_peak")
        return tuple(out)

    def margincalc(self, tid, param, x, logi=1, yadjust=0.0, expo=False, mean=1.0):
        x = np.ascontiguousarray(np.atleast_1d(x), np.float64)
        out = np.zeros(len(x))
        self._chk(self.lib.ima_lmode_margincalc(self.h, tid, param, x.ctypes.data_as(L.dp),
                                                len(x), int(logi), float(yadjust), int(expo),
                                                float(mean), out.ctypes.data_as(L.dp)),
                  "margincalc")
        return out

    def popmig(self, tid, theta_i, mig_i, x, qmax, mprior, expo=False, prob_or_like=0,
               kind=L.POPMIG_CALC, first=0, last=0):
        """2NM density at x (ima_lmode_popmig): calc_popmig / calc_pop_expomig,
        or with kind=POPMIG_MARGIN marginpopmig over rows [first, last)."""
        x = np.ascontiguousarray(np.atleast_1d(x), np.float64)
        out = np.zeros(len(x))
        self._chk(self.lib.ima_lmode_popmig(self.h, int(tid), int(theta_i), int(mig_i),
                                            float(qmax), float(mprior), int(expo), int(kind),
                                            int(prob_or_like), int(first), int(last),
                                            x.ctypes.data_as(L.dp), len(x),
                                            out.ctypes.data_as(L.dp)), "popmig")
        return out

    def popmig_peak(self, tid, theta_i, mig_i, qmax, mprior, expo=False, first=0, last=None):
        """(peak, density there, points evaluated) of the 2NM density
        (ima_lmode_popmig_peak, marginalopt_popmig for one pair)."""
        if last is None:
            last = self.lmode_table_rows(tid)[0]
        pk, ml, ev = C.c_double(), C.c_double(), C.c_int64()
        self._chk(self.lib.ima_lmode_popmig_peak(self.h, int(tid), int(theta_i), int(mig_i),
                                                 float(qmax), float(mprior), int(expo),
                                                 int(first), int(last), C.byref(pk),
                                                 C.byref(ml), C.byref(ev)), "popmig_peak")
        return pk.value, ml.value, ev.value

    def greater_than(self, tid, kind, prior_max):
        """P(row parameter > column parameter) for the population sizes
        (kind 0) or migration rates (kind 1) with the given prior maxima
        (ima_lmode_greater_than); -1 where the reference prints "na"."""
        pm = np.ascontiguousarray(prior_max, np.float64)
        n = len(pm)
        out = np.zeros((n, n))
        self._chk(self.lib.ima_lmode_greater_than(self.h, int(tid), int(kind), n,
                                                  pm.ctypes.data_as(L.dp),
                                                  out.ctypes.data_as(L.dp)), "greater_than")
        return out

    # --------------------------------------------------------------- MC3 --
    def mc3_setup(self, nchains, nloci, first_slot, u, kappa, beta, seed=1,
                  precision=L.PREC_FP32, swaptries=None, scale_width=0.2):
        u, pu = _f64(np.broadcast_to(np.asarray(u, float), (nchains * nloci,)))
        k, pk = _f64(np.broadcast_to(np.asarray(kappa, float), (nchains * nloci,)))
        b, pb = _f64(beta)
        if swaptries is None:
            swaptries = max(1, nchains // 10)  # ima_main_mpi.cpp:1376-1378
        self._chk(self.lib.ima_mc3_setup(self.h, nchains, nloci, first_slot, pu, pk, pb,
                                         int(seed), precision, int(swaptries),
                                         float(scale_width)), "mc3_setup")
        self.mc3_shape = (nchains, nloci)

    def mc3_run(self, nsweeps, flags=L.MC3_GRAPH | L.MC3_SCALE):
        self._chk(self.lib.ima_mc3_run(self.h, int(nsweeps), int(flags)), "mc3_run")

    def mc3_read(self):
        C_, L_ = self.mc3_shape
        pdg = np.zeros(C_ * L_)
        beta = np.zeros(C_)
        stats = np.zeros(8, np.uint64)
        sweeps = C.c_uint64()
        self._chk(self.lib.ima_mc3_read(self.h, pdg.ctypes.data_as(L.dp),
                                        beta.ctypes.data_as(L.dp),
                                        stats.ctypes.data_as(C.POINTER(C.c_uint64)),
                                        C.byref(sweeps)), "mc3_read")
        return pdg.reshape(C_, L_), beta, stats, sweeps.value

    def mc3_decide(self, prop_first, hastings=None, uniforms=None, gbeta=1.0):
        """Decisions for host-written proposals in slots prop_first + c*L + l
        (ima_mc3_decide); returns the accepted flags [nchains, nloci]."""
        C_, L_ = self.mc3_shape
        acc = np.zeros(C_ * L_, np.int32)
        h = None if hastings is None else np.ascontiguousarray(np.ravel(hastings), np.float64)
        u = None if uniforms is None else np.ascontiguousarray(np.ravel(uniforms), np.float64)
        self._chk(self.lib.ima_mc3_decide(
            self.h, int(prop_first), None if h is None else h.ctypes.data_as(L.dp),
            None if u is None else u.ctypes.data_as(L.dp), float(gbeta),
            acc.ctypes.data_as(L.i32p)), "mc3_decide")
        return acc.reshape(C_, L_)

    def mc3_set_prior(self, tvals, tmax, P):
        """Genealogy prior in the sweep; tvals [nchains, P-1] split times."""
        C_, _ = self.mc3_shape
        tv = np.ascontiguousarray(np.asarray(tvals, float).reshape(C_, P - 1) if P > 1
                                  else np.zeros((C_, 1)))
        self.mc3_P = P
        self._chk(self.lib.ima_mc3_set_prior(self.h, tv.ctypes.data_as(L.dp), float(tmax)),
                  "mc3_set_prior")

    def mc3_set_analyses(self, chains_per_analysis):
        """Split the chains into uncoupled analyses of `chains_per_analysis`
        consecutive chains each: swaps stay inside an analysis."""
        self._chk(self.lib.ima_mc3_set_analyses(self.h, int(chains_per_analysis)),
                  "mc3_set_analyses")

    def mc3_prior_group(self, loci=None):
        """Loci per speculative group of the accept step; sets it when
        `loci` is given.  Returns the value in force."""
        if loci is not None:
            self._chk(self.lib.ima_mc3_set_prior_group(self.h, int(loci)), "mc3_set_prior_group")
        out = C.c_int()
        self._chk(self.lib.ima_mc3_prior_group(self.h, C.byref(out)), "mc3_prior_group")
        return out.value

    def mc3_tune(self, sweeps=20, flags=None):
        """Autotune the accept step's group size on the sweeps `flags`
        describe (ima_mc3_run's flags; default GRAPH | SCALE); the chains
        are left as they were.  Returns (chosen size, ms per sweep for
        sizes 1..6)."""
        if flags is None:
            flags = L.MC3_GRAPH | L.MC3_SCALE
        best = C.c_int()
        ms = np.zeros(6)
        self._chk(self.lib.ima_mc3_tune(self.h, int(sweeps), int(flags), C.byref(best),
                                        ms.ctypes.data_as(L.dp)), "mc3_tune")
        return best.value, ms

    def set_site_block(self, patterns):
        """HKY site-block size: 0 automatic, 64 or 128 patterns per unit."""
        self._chk(self.lib.ima_set_site_block(self.h, int(patterns)), "set_site_block")

    def mc3_site_block(self):
        out = C.c_int()
        self._chk(self.lib.ima_mc3_site_block(self.h, C.byref(out)), "mc3_site_block")
        return out.value

    def mc3_read_rates(self):
        """Each chain's mutation-rate scalars and kappa, [nchains, nloci] each."""
        C_, L_ = self.mc3_shape
        u = np.zeros(C_ * L_)
        k = np.zeros(C_ * L_)
        self._chk(self.lib.ima_mc3_read_rates(self.h, u.ctypes.data_as(L.dp),
                                              k.ctypes.data_as(L.dp)), "mc3_read_rates")
        return u.reshape(C_, L_), k.reshape(C_, L_)

    def mc3_read_prior(self):
        C_, _ = self.mc3_shape
        P = self.mc3_P
        probg = np.zeros(C_)
        tv = np.zeros((C_, max(P - 1, 1)))
        self._chk(self.lib.ima_mc3_read_prior(self.h, probg.ctypes.data_as(L.dp),
                                              tv.ctypes.data_as(L.dp)), "mc3_read_prior")
        return probg, tv[:, :P - 1]

    def read_mig_events(self, slot):
        """(edge, time, to) arrays of one slot's migration events."""
        cnt = C.c_int32()
        self._chk(self.lib.ima_slots_read_mig_events(self.h, slot, C.byref(cnt), None, None,
                                                     None), "read_mig_events")
        m = cnt.value
        e, t, to = np.zeros(max(m, 1), np.int32), np.zeros(max(m, 1)), np.zeros(max(m, 1),
                                                                             np.int32)
        if m:
            self._chk(self.lib.ima_slots_read_mig_events(
                self.h, slot, C.byref(cnt), e.ctypes.data_as(L.i32p), t.ctypes.data_as(L.dp),
                to.ctypes.data_as(L.i32p)), "read_mig_events")
        return e[:m], t[:m], to[:m]

    def mcf_save(self, path, first, nchains, nloci, layout, pi, u=None, kappa=None,
                 tvals=None):
        """Write the chains' genealogies (slots first.., chain after chain)
        and rates / split times to a .mcf file (ima_mcf_write, writemcf).
        layout: [(model, n, S)] per locus; pi [nloci, 4] base frequencies.
        u / kappa / tvals default to the MC3 state's."""
        from . import mcffile
        N = nchains * nloci
        up0, up1, dn, tm, rt, A = self.read_slots(first, N)
        pop = self.read_pops(first, N)
        counts, mt, mo = [], [], []
        for i in range(N):
            n = layout[i % nloci][1]
            e, t, to = self.read_mig_events(first + i)
            c = np.bincount(e, minlength=2 * n - 1) if len(e) else np.zeros(2 * n - 1, int)
            counts.append(c)
            order = np.lexsort((t, e))
            mt.append(t[order])
            mo.append(to[order])
        if u is None or kappa is None:
            ru, rk = self.mc3_read_rates()
            u = ru if u is None else u
            kappa = rk if kappa is None else kappa
        if tvals is None:
            tvals = self.mc3_read_prior()[1]
        tvals = np.asarray(tvals, float).reshape(nchains, -1)
        state = {"tvals": tvals, "u": np.asarray(u, float).reshape(nchains, nloci),
                 "kappa": np.asarray(kappa, float).reshape(nchains, nloci),
                 "pi": np.tile(np.asarray(pi, float).reshape(1, nloci, 4), (nchains, 1, 1)),
                 "up0": up0, "up1": up1, "down": dn, "time": tm, "pop": pop, "A": A,
                 "mig_count": np.concatenate(counts).astype(np.int32),
                 "mig_time": np.concatenate(mt) if mt else np.zeros(0),
                 "mig_to": np.concatenate(mo).astype(np.int32) if mo else np.zeros(0, np.int32)}
        mcffile.write(path, nchains, tvals.shape[1], layout, state)

    def mcf_load(self, path, first, nchains, nloci, layout, tmax=None):
        """Read a .mcf file (ima_mcf_read, readmcf) into slots first..
        (genealogies, populations, migrations); returns the state dict
        (tvals, u, kappa, pi, ...) for mc3_setup / mc3_set_prior, which
        re-score everything from the genealogies as readmcf's init_p does."""
        from . import mcffile
        from .sim import Genealogy
        nsplit = 0 if tmax is None else len(tmax)
        st = mcffile.read(path, nchains, nsplit, layout, tmax=tmax) if tmax is not None else \
            None
        if st is None:
            raise ValueError("mcf_load needs tmax (one bound per split time)")
        N = nchains * nloci
        o = 0
        gens, A, moff = [], [], 0
        for i in range(N):
            n = layout[i % nloci][1]
            E = 2 * n - 1
            sl = slice(o, o + E)
            dn = st["down"][sl]
            g = Genealogy(n=n, up0=st["up0"][sl].copy(), up1=st["up1"][sl].copy(), down=dn.copy(),
                          time=st["time"][sl].copy(), root=int(np.flatnonzero(dn == -1)[0]),
                          pop=st["pop"][sl].copy())
            cnt = st["mig_count"][sl]
            mig = [[] for _ in range(E)]
            for e in range(E):
                for _ in range(int(cnt[e])):
                    mig[e].append((float(st["mig_time"][moff]), int(st["mig_to"][moff])))
                    moff += 1
            g.mig = mig
            gens.append(g)
            A.append(st["A"][sl].copy())
            o += E
        self.write_slots(first, gens, A=A)
        self.write_pops(first, gens)
        return st

    def read_mig_times(self, slot, nmig):
        out = np.zeros(max(nmig, 1))
        self._chk(self.lib.ima_slots_read_mig(self.h, slot, out.ctypes.data_as(L.dp)),
                  "read_mig_times")
        return out[:nmig]

    # ------------------------------------------------------------ timing --
    def timing(self, on=True):
        self._chk(self.lib.ima_timing_enable(self.h, int(on)), "timing")

    def timing_get(self, k):
        n = C.c_int64()
        ms = C.c_double()
        self._chk(self.lib.ima_timing_get(self.h, k, C.byref(n), C.byref(ms)), "timing_get")
        return n.value, ms.value

    def timing_reset(self):
        self._chk(self.lib.ima_timing_reset(self.h), "timing_reset")


class RefShim:
    """The reference's per-(chain, locus) calls over engine slots.

    ``trees[ci][li]`` holds the Genealogy of chain ci at locus li.  Each
    call uploads that genealogy to its slot and scores it, like the
    reference's in-place evaluation of C[ci]->G[li].
    """

    def __init__(self, engine: Engine, loci_ids, nchains):
        self.e = engine
        self.loci_ids = list(loci_ids)
        self.nchains = nchains
        slots = [l for _ in range(nchains) for l in self.loci_ids]
        self.first = engine.alloc_slots(slots)
        self.sumlogk_done = set()

    def slot(self, ci, li):
        return self.first + ci * len(self.loci_ids) + li

    def likelihoodHKY(self, ci, li, g, mutrate, kappa, e1=-1, e2=-1, e3=-1, e4=-1):
        s = self.slot(ci, li)
        self.e.write_slots(s, [g])
        out, _ = self.e.score(self.e.items([s], mutrate, kappa))
        return float(out[0])

    def likelihoodIS(self, ci, li, g, mutrate):
        s = self.slot(ci, li)
        self.e.write_slots(s, [g])
        if li not in self.sumlogk_done:  # first genealogy seen (calc_prob_data.cpp:760-765)
            self.e.sumlogk_from_slot(self.loci_ids[li], s)
            self.sumlogk_done.add(li)
        out, _ = self.e.score(self.e.items([s], mutrate))
        return float(out[0])

    def likelihoodSW(self, ci, li, g, A, u, tr=1.0):
        s = self.slot(ci, li)
        self.e.write_slots(s, [g], A=[A])
        out, _ = self.e.score(self.e.items([s], u, tr=tr))
        return float(out[0])
