"""ctypes front-end of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
RNG_REFERENCE, RNG_COUNTER = 0, 1


class PanelDesc(ctypes.Structure):
    _fields_ = [("n_sites", ctypes.c_int32), ("n_haps", ctypes.c_int32),
                ("cm", ctypes.c_void_p), ("cref", ctypes.c_void_p), ("calt", ctypes.c_void_p),
                ("lq", ctypes.c_void_p), ("common_words", ctypes.c_void_p),
                ("rare_ptr", ctypes.c_void_p), ("rare_idx", ctypes.c_void_p)]


class Params(ctypes.Structure):
    _fields_ = [("burnin", ctypes.c_int32), ("main_iter", ctypes.c_int32), ("ne", ctypes.c_int32),
                ("min_gl", ctypes.c_float), ("err_imp", ctypes.c_float), ("err_phase", ctypes.c_float),
                ("pbwt_depth", ctypes.c_int32), ("pbwt_modulo_cm", ctypes.c_float),
                ("kinit", ctypes.c_int32), ("kpbwt", ctypes.c_int32), ("sparse_maf", ctypes.c_float),
                ("seed", ctypes.c_uint32), ("rng_mode", ctypes.c_int32), ("n_threads", ctypes.c_int32),
                ("ploidy", ctypes.c_void_p)]


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_panel_create.restype = vp
        L.orc_panel_create.argtypes = [ctypes.POINTER(PanelDesc), ctypes.c_float]
        L.orc_panel_free.argtypes = [vp]
        L.orc_panel_n_common.argtypes = [vp]
        L.orc_panel_n_common_hq.argtypes = [vp]
        L.orc_panel_site_class.argtypes = [vp, vp, vp]
        L.orc_panel_ypacked.restype = ctypes.c_int64
        L.orc_panel_ypacked.argtypes = [vp, vp, ctypes.c_int64]
        L.orc_panel_asmall.restype = ctypes.c_int64
        L.orc_panel_asmall.argtypes = [vp, vp, vp, ctypes.c_int64]
        L.orc_default_params.argtypes = [ctypes.POINTER(Params)]
        L.orc_phase_run.argtypes = [vp, ctypes.c_int32, vp, vp, ctypes.POINTER(Params), vp,
                                    vp, vp, vp, vp, vp, vp]
        L.orc_pbwt_select.argtypes = [vp, ctypes.c_int32, vp, vp, vp, ctypes.c_int32, ctypes.c_float,
                                      vp, vp, vp, ctypes.c_int64]
        L.orc_pbwt_trace.argtypes = [vp, ctypes.c_int32, vp, vp, vp, ctypes.c_int32, ctypes.c_float, vp,
                                     ctypes.c_int64, vp]
        L.orc_pbwt_groups.argtypes = [vp, ctypes.c_int32, ctypes.c_float, ctypes.c_int32, vp, vp,
                                      ctypes.c_int32, vp]
        L.orc_condset.argtypes = [vp, ctypes.c_int32, vp, vp, vp, ctypes.c_int32, ctypes.c_int32,
                                  vp, vp, vp, vp, vp, ctypes.c_int64, vp]
        L.orc_condset_from_states.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32, vp, vp, vp,
                                              ctypes.c_int64, vp]
        L.orc_impute_hap.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, vp, vp, vp]
        L.orc_hap_likelihoods.argtypes = [ctypes.c_int32, vp, vp, vp, ctypes.c_int32, ctypes.c_float, vp]
        L.orc_rephase.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, vp, vp, vp,
                                  ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32]
        L.orc_phase_trace.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, vp, vp, vp,
                                      ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      vp, vp, vp, vp, vp]
        L.orc_philox4x32.argtypes = [vp, vp, vp]
        L.orc_mt19937_nth.restype = ctypes.c_uint32
        L.orc_mt19937_nth.argtypes = [ctypes.c_uint32, ctypes.c_int64]
        L.orc_unphred.restype = ctypes.c_float
        L.orc_unphred.argtypes = [ctypes.c_int32]
        L.orc_output_fields.argtypes = [ctypes.c_int32, ctypes.c_int32] + [ctypes.c_void_p] * 11
        L.orc_output_fields_ploidy.argtypes = [ctypes.c_int32, ctypes.c_int32] + [ctypes.c_void_p] * 12
        i32, u32 = ctypes.c_int32, ctypes.c_uint32
        L.orc_sample.argtypes = [i32, i32, i32, u32, i32, vp, vp]
        L.orc_init_states.argtypes = [vp, vp, vp, i32, i32, u32, i32, vp, i32, vp]
        L.orc_gp_run.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp, vp]
        L.orc_draws.argtypes = [u32, i32, i32, i32, i32, i32, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def default_params(**kw):
    p = Params()
    lib().orc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Panel:
    """Oracle-side panel (classification + sparse PBWT build)."""

    def __init__(self, P, sparse_maf=0.001):
        self._keep = dict(cm=np.ascontiguousarray(P["cm"], np.float64),
                          cref=np.ascontiguousarray(P["cref"], np.uint32),
                          calt=np.ascontiguousarray(P["calt"], np.uint32),
                          lq=np.ascontiguousarray(P.get("lq", np.zeros(P["L"], np.uint8)), np.uint8),
                          cw=np.ascontiguousarray(P["common_words"], np.uint64),
                          rp=np.ascontiguousarray(P["rare_ptr"], np.int32),
                          ri=np.ascontiguousarray(P["rare_idx"], np.int32))
        k = self._keep
        d = PanelDesc(P["L"], P["N"], _p(k["cm"]).value, _p(k["cref"]).value, _p(k["calt"]).value,
                      _p(k["lq"]).value, _p(k["cw"]).value, _p(k["rp"]).value, _p(k["ri"]).value)
        self._free = lib().orc_panel_free   # held so that teardown at interpreter exit can still free
        self.h = lib().orc_panel_create(ctypes.byref(d), ctypes.c_float(sparse_maf))
        if not self.h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())
        self.L, self.N = P["L"], P["N"]

    def __del__(self):
        if getattr(self, "h", None):
            self._free(self.h)
            self.h = None

    @property
    def n_common(self):
        return lib().orc_panel_n_common(self.h)

    @property
    def n_common_hq(self):
        return lib().orc_panel_n_common_hq(self.h)

    def site_class(self):
        fc = np.zeros(self.L, np.uint8)
        mj = np.zeros(self.L, np.uint8)
        lib().orc_panel_site_class(self.h, _p(fc), _p(mj))
        return fc, mj

    def ypacked(self):
        n = lib().orc_panel_ypacked(self.h, None, 0)
        out = np.zeros(n, np.uint8)
        lib().orc_panel_ypacked(self.h, _p(out), n)
        return out

    def asmall(self):
        n = lib().orc_panel_asmall(self.h, None, None, 0)
        ptr = np.zeros(self.n_common_hq + 2, np.int32)
        idx = np.zeros(max(n, 1), np.int32)
        lib().orc_panel_asmall(self.h, _p(ptr), _p(idx), max(n, 1))
        return ptr, idx[:n]

    def groups(self, depth, modulo, kpbwt):
        eff = ctypes.c_float()
        n = ctypes.c_int32()
        grp = np.zeros(self.L + 2, np.int32)
        _check(lib().orc_pbwt_groups(self.h, depth, ctypes.c_float(modulo), kpbwt, ctypes.byref(eff),
                                     _p(grp), grp.size, ctypes.byref(n)))
        return eff.value, grp[:n.value]

    def pbwt_trace(self, hap_alleles, stored, start_idx, depth=12, modulo=0.1):
        """Rows (site, haplotype, idx, f, g, last reset of [f, g) or -1) after every
        HQ common site of the sweep, per target haplotype."""
        haps = np.ascontiguousarray(hap_alleles, np.uint8)
        T = haps.shape[0] // 2
        cap = int(haps.shape[0]) * int(haps.shape[1]) + 1
        rows = np.zeros((cap, 6), np.int32)
        n = ctypes.c_int64()
        _check(lib().orc_pbwt_trace(self.h, T, _p(haps), _p(np.ascontiguousarray(stored, np.uint8)),
                                    _p(np.ascontiguousarray(start_idx, np.int32)), depth, ctypes.c_float(modulo),
                                    _p(rows), cap, ctypes.byref(n)))
        return rows[:n.value]

    def pbwt_select(self, hap_alleles, stored, start_idx, depth=12, modulo=0.1, cap=None):
        haps = np.ascontiguousarray(hap_alleles, np.uint8)
        T = haps.shape[0] // 2
        cap = cap or max(1 << 16, 64 * depth * T * 64)
        off = np.zeros(T * depth, np.int32)
        cnt = np.zeros(T * depth, np.int32)
        ids = np.zeros(cap, np.int32)
        _check(lib().orc_pbwt_select(self.h, T, _p(haps), _p(np.ascontiguousarray(stored, np.uint8)),
                                     _p(np.ascontiguousarray(start_idx, np.int32)), depth,
                                     ctypes.c_float(modulo), _p(off), _p(cnt), _p(ids), cap))
        lists = []
        for t in range(T):
            lists.append([ids[off[t * depth + d]: off[t * depth + d] + cnt[t * depth + d]].copy()
                          for d in range(depth)])
        return lists

    def condset(self, lists, kpbwt, ne=100000):
        depth = len(lists)
        off = np.zeros(depth, np.int32)
        cnt = np.array([len(x) for x in lists], np.int32)
        off[1:] = np.cumsum(cnt)[:-1]
        ids = np.concatenate([np.asarray(x, np.int32) for x in lists] + [np.zeros(1, np.int32)])
        states = np.zeros(self.N, np.int32)
        K = ctypes.c_int32()
        poly = np.zeros(self.L, np.int32)
        npoly = ctypes.c_int32()
        cap = self.L * ((self.N + 7) // 8)
        hv = np.zeros(cap, np.uint8)
        t = np.zeros(self.L, np.float32)
        _check(lib().orc_condset(self.h, depth, _p(off), _p(cnt), _p(ids), kpbwt, ne, _p(states),
                                 ctypes.byref(K), _p(poly), ctypes.byref(npoly), _p(hv), cap, _p(t)))
        K, n = K.value, npoly.value
        stride = (K + 7) // 8
        return dict(states=states[:K], poly=poly[:n], hvar=hv[:n * stride].reshape(n, stride),
                    t=t[:max(n - 1, 0)])

    def condset_from_states(self, states, ne=100000):
        st = np.ascontiguousarray(states, np.int32)
        K = st.size
        poly = np.zeros(self.L, np.int32)
        npoly = ctypes.c_int32()
        stride = (K + 7) // 8
        cap = self.L * stride
        hv = np.zeros(max(cap, 1), np.uint8)
        t = np.zeros(self.L, np.float32)
        _check(lib().orc_condset_from_states(self.h, _p(st), K, ne, _p(poly), ctypes.byref(npoly),
                                             _p(hv), cap, _p(t)))
        n = npoly.value
        return dict(states=st, poly=poly[:n], hvar=hv[:n * stride].reshape(n, stride),
                    t=t[:max(n - 1, 0)])

    def impute_hap(self, states, HL, flat, ne=100000, err_imp=1e-12):
        st = np.ascontiguousarray(states, np.int32)
        HP = np.zeros(2 * self.L, np.float32)
        _check(lib().orc_impute_hap(self.h, _p(st), st.size, ne, ctypes.c_float(err_imp),
                                    _p(np.ascontiguousarray(HL, np.float32)),
                                    _p(np.ascontiguousarray(flat, np.uint8)), _p(HP)))
        return HP

    def rephase(self, states, flat, h0, h1, seed, it, tgt, ne=100000, err_phase=1e-4):
        st = np.ascontiguousarray(states, np.int32)
        a = np.ascontiguousarray(h0, np.uint8).copy()
        b = np.ascontiguousarray(h1, np.uint8).copy()
        _check(lib().orc_rephase(self.h, _p(st), st.size, ne, ctypes.c_float(err_phase),
                                 _p(np.ascontiguousarray(flat, np.uint8)), _p(a), _p(b), seed, it, tgt))
        return a, b

    def phase_trace(self, states, flat, h0, h1, seed=1, it=1, tgt=0, ne=100000, err_phase=1e-4, cap=4096):
        """The deterministic quantities of one rephase: per segment start the
        backward sums per hypothesis over their total [n_segs][8], per segment
        boundary TRANS_HAP's joint hypothesis probabilities [n_segs-1][8][8], per
        flat het P(allele 1 | hypothesis) [n_miss][8]."""
        st = np.ascontiguousarray(states, np.int32)
        ns, nm = ctypes.c_int32(), ctypes.c_int32()
        seg = np.zeros((cap, 8), np.float32)
        hp = np.zeros((cap, 8, 8), np.float32)
        of1 = np.zeros((cap, 8), np.float32)
        _check(lib().orc_phase_trace(self.h, _p(st), st.size, ne, ctypes.c_float(err_phase),
                                     _p(np.ascontiguousarray(flat, np.uint8)),
                                     _p(np.ascontiguousarray(h0, np.uint8)), _p(np.ascontiguousarray(h1, np.uint8)),
                                     seed, it, tgt, cap, cap, ctypes.byref(ns), ctypes.byref(nm), _p(seg), _p(hp),
                                     _p(of1)))
        n, m = ns.value, nm.value
        return seg[:n], hp[:max(n - 1, 0)], of1[:m]

    def init_states(self, gl, flat, kinit, rng_mode, seed, tgt=0):
        out = np.zeros(self.N, np.int32)
        n = ctypes.c_int32()
        _check(lib().orc_init_states(self.h, _p(np.ascontiguousarray(gl, np.uint8)),
                                     _p(np.ascontiguousarray(flat, np.uint8)), kinit, rng_mode, seed, tgt, _p(out),
                                     self.N, ctypes.byref(n)))
        return out[:n.value]

    def phase_run(self, gl, flat, params, kpbwt_per_target=None, ploidy=None):
        gl = np.ascontiguousarray(gl, np.uint8)
        flat = np.ascontiguousarray(flat, np.uint8)
        T, L = flat.shape
        n_iter = 1 + params.burnin + params.main_iter
        gp0 = np.zeros((T, L), np.float32)
        gp1 = np.zeros((T, L), np.float32)
        hds = np.zeros((T, L), np.uint8)
        h0 = np.zeros((T, L), np.uint8)
        h1 = np.zeros((T, L), np.uint8)
        ns = np.zeros((T, n_iter), np.int32)
        kpt = None if kpbwt_per_target is None else np.ascontiguousarray(kpbwt_per_target, np.int32)
        pl = None
        if ploidy is not None:
            pl = np.ascontiguousarray(np.broadcast_to(np.asarray(ploidy, np.int32), (T,)), np.int32)
            params.ploidy = pl.ctypes.data
        else:
            params.ploidy = None
        _check(lib().orc_phase_run(self.h, T, _p(gl), _p(flat), ctypes.byref(params), _p(kpt),
                                   _p(gp0), _p(gp1), _p(hds), _p(h0), _p(h1), _p(ns)))
        return dict(gp0=gp0, gp1=gp1, hds=hds, h0=h0, h1=h1, nstates=ns)


def hap_likelihoods(gl, flat, other=None, min_gl=1e-10, HL=None):
    gl = np.ascontiguousarray(gl, np.uint8)
    L = gl.shape[0]
    HL = np.zeros(2 * L, np.float32) if HL is None else np.ascontiguousarray(HL, np.float32).copy()
    mode = 0 if other is None else 1
    oth = np.zeros(L, np.uint8) if other is None else np.ascontiguousarray(other, np.uint8)
    _check(lib().orc_hap_likelihoods(L, _p(gl), _p(np.ascontiguousarray(flat, np.uint8)), _p(oth), mode,
                                     ctypes.c_float(min_gl), _p(HL)))
    return HL


def philox4x32(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox4x32(_p(c), _p(k), _p(o))
    return o


def mt19937_nth(seed, n):
    return lib().orc_mt19937_nth(seed, n)


def unphred(i):
    return lib().orc_unphred(i)


def output_fields(gp0, gp1, hds, cref, calt, ploidy=None):
    """genotype_writer per-site fields (DS, GP, GT, INFO, AF, RAF); ploidy [T] (None: diploid)."""
    gp0 = np.ascontiguousarray(gp0, np.float32)
    gp1 = np.ascontiguousarray(gp1, np.float32)
    hds = np.ascontiguousarray(hds, np.uint8)
    cref = np.ascontiguousarray(cref, np.uint32)
    calt = np.ascontiguousarray(calt, np.uint32)
    T, L = gp0.shape
    out = dict(ds=np.zeros((T, L), np.float32), gp=np.zeros((T, L, 3), np.float32), gt=np.zeros((T, L, 2), np.uint8),
               info=np.zeros(L, np.float32), af=np.zeros(L, np.float32), raf=np.zeros(L, np.float32))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    pl = None if ploidy is None else np.ascontiguousarray(np.broadcast_to(np.asarray(ploidy, np.int32), (T,)), np.int32)
    rc = lib().orc_output_fields_ploidy(T, L, None if pl is None else p(pl), p(gp0), p(gp1), p(hds), p(cref), p(calt),
                                        p(out["ds"]), p(out["gp"]), p(out["gt"]), p(out["info"]), p(out["af"]),
                                        p(out["raf"]))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return out


def sample(n_pop, k, rng_mode, seed, tgt=0):
    """One INIT sampling call over 0..n_pop-1 (std::sample or Algorithm S)."""
    out = np.zeros(max(k, 1), np.int32)
    n = ctypes.c_int32()
    _check(lib().orc_sample(n_pop, k, rng_mode, seed, tgt, _p(out), ctypes.byref(n)))
    return out[:n.value]


def draws(seed, it, purpose, a, n, rng_mode):
    out = np.zeros(n, np.float32)
    _check(lib().orc_draws(seed, it, purpose, a, n, rng_mode, _p(out)))
    return out


def gp_run(HP0, HP1):
    """GP storage over the main iterations of one target: HP0/HP1 [n_iter][2L]."""
    HP0 = np.ascontiguousarray(HP0, np.float32)
    HP1 = np.ascontiguousarray(HP1, np.float32)
    n_iter, L2 = HP0.shape
    L = L2 // 2
    out = dict(gp0=np.zeros(L, np.float32), gp1=np.zeros(L, np.float32), hds=np.zeros(L, np.uint8),
               h0=np.zeros(L, np.uint8), h1=np.zeros(L, np.uint8))
    _check(lib().orc_gp_run(L, n_iter, _p(HP0), _p(HP1), _p(out["gp0"]), _p(out["gp1"]), _p(out["hds"]),
                            _p(out["h0"]), _p(out["h1"])))
    return out
