"""ctypes front-end of libglimpse_hip.so (the C-ABI in include/glimpse_hip.h).

This is the Python host mirror of the phase module's compute interfaces:
Panel ~ the loaded split_reference panel, condset_batch ~ conditioning_set::select,
impute_batch ~ imputation_hmm::computePosteriors, rephase_batch ~
phasing_hmm::rephaseHaplotypes, pbwt_select ~
haplotype_set::matchHapsFromCompressedPBWTSmall, phase_run ~ caller::phase_loop.
There is no CPU fallback: every call goes to the HIP library and raises if it
is missing or reports an error."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libglimpse_hip.so")
_lib = None
vp = ctypes.c_void_p


class GhError(RuntimeError):
    pass


class PanelDesc(ctypes.Structure):
    _fields_ = [("n_sites", ctypes.c_int32), ("n_haps", ctypes.c_int32), ("cm", vp), ("cref", vp),
                ("calt", vp), ("lq", vp), ("common_words", vp), ("rare_ptr", vp), ("rare_idx", vp),
                ("ypacked", vp), ("ypacked_len", ctypes.c_int64), ("asmall_ptr", vp), ("asmall_idx", vp),
                ("sparse_maf", ctypes.c_float)]


class Params(ctypes.Structure):
    _fields_ = [("burnin", ctypes.c_int32), ("main_iter", ctypes.c_int32), ("ne", ctypes.c_int32),
                ("min_gl", ctypes.c_float), ("err_imp", ctypes.c_float), ("err_phase", ctypes.c_float),
                ("pbwt_depth", ctypes.c_int32), ("pbwt_modulo_cm", ctypes.c_float),
                ("kinit", ctypes.c_int32), ("kpbwt", ctypes.c_int32), ("seed", ctypes.c_uint32)]


class RunStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "fb_state_sites", "fb_poly_sites", "ph_state_sites", "ph_segments", "pbwt_hap_sites",
        "ms_select", "ms_condset", "ms_fb", "ms_phase", "ms_other", "ms_total",
        "fb_kernel_ms", "fb_launches", "fb_bytes", "ms_learn", "learned_sel", "pool_kept", "collected_rows",
        "fb_nw", "fb_mwl", "ph_nw", "ph_kpt", "graph_iters", "graph_reruns", "ms_graph")]

MLP_IN, MLP_HIDDEN, MLP_NPARAM = 32, 128, 20865


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GhError("libglimpse_hip.so is not built; run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        i32, i64, f32, u32 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_uint32
        L.gh_last_error.restype = ctypes.c_char_p
        L.gh_default_params.argtypes = [ctypes.POINTER(Params)]
        L.gh_build_sparse_pbwt.argtypes = [ctypes.POINTER(PanelDesc), vp, vp, vp, vp, vp, vp]
        L.gh_panel_create.argtypes = [ctypes.POINTER(PanelDesc), ctypes.POINTER(vp)]
        L.gh_panel_destroy.argtypes = [vp]
        L.gh_panel_info.argtypes = [vp, vp, vp, vp]
        L.gh_condset_batch.argtypes = [vp, i32, vp, vp, i32, vp, vp, i64, vp, vp, i64, vp]
        L.gh_impute_batch.argtypes = [vp, i32, vp, vp, vp, vp, f32, i32, vp]
        L.gh_impute_batch_noalloc.argtypes = [vp, i32, vp, vp, vp, vp, f32, i32, vp]
        L.gh_rephase_batch.argtypes = [vp, i32, vp, vp, vp, vp, vp, f32, i32, u32, i32, vp]
        L.gh_pbwt_select.argtypes = [vp, i32, vp, vp, vp, i32, f32, vp, vp, vp, i64, i32, vp, vp]
        L.gh_phase_run.argtypes = [vp, i32, vp, vp, ctypes.POINTER(Params), vp, vp, vp, vp, vp, vp, vp,
                                   ctypes.POINTER(RunStats)]
        L.gh_run_create.argtypes = [vp, i32, vp, vp, ctypes.POINTER(Params), vp, ctypes.POINTER(vp)]
        L.gh_run_execute.argtypes = [vp]
        L.gh_run_fetch.argtypes = [vp, vp, vp, vp, vp, vp, vp, ctypes.POINTER(RunStats)]
        L.gh_run_destroy.argtypes = [vp]
        L.gh_workspace_create.argtypes = [ctypes.POINTER(vp)]
        L.gh_workspace_destroy.argtypes = [vp]
        L.gh_workspace_bytes.argtypes = [vp, vp]
        L.gh_run_set_workspace.argtypes = [vp, vp]
        L.gh_mlp_create.argtypes = [u32, ctypes.POINTER(vp)]
        L.gh_mlp_destroy.argtypes = [vp]
        L.gh_mlp_get_params.argtypes = [vp, vp]
        L.gh_mlp_set_params.argtypes = [vp, vp]
        L.gh_mlp_forward.argtypes = [vp, i64, vp, vp]
        L.gh_mlp_backward.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        L.gh_mlp_step.argtypes = [vp, i32, vp, vp, vp, f32, vp]
        L.gh_mlp_train_run.argtypes = [vp, vp, i32, i32, f32, u32, vp]
        L.gh_run_collect.argtypes = [vp, i32]
        L.gh_run_set_tuning.argtypes = [vp, i32]
        L.gh_run_set_graph.argtypes = [vp, i32]
        L.gh_init_states.argtypes = [vp, i32, vp, vp, i32, u32, i32, vp, vp]
        L.gh_run_set_selector_prior.argtypes = [vp, f32]
        L.gh_run_collected.argtypes = [vp, vp, vp, vp, vp]
        L.gh_run_set_selector.argtypes = [vp, vp, vp, f32, i32]
        L.gh_select_features.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, i32, i32, vp]
        L.gh_select_topk.argtypes = [i32, vp, vp, vp, vp, i32, f32, vp, vp, vp]
        L.gh_output_fields.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.gh_output_fields_ploidy.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.gh_run_set_ploidy.argtypes = [vp, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(vp)


def _check(rc):
    if rc != 0:
        raise GhError("glimpse_hip error %d: %s" % (rc, lib().gh_last_error().decode()))


def default_params(**kw):
    p = Params()
    lib().gh_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _csr(lists):
    lists = [np.asarray(x, np.int32) for x in lists]
    ptr = np.zeros(len(lists) + 1, np.int32)
    ptr[1:] = np.cumsum([x.size for x in lists])
    ids = np.concatenate(lists + [np.zeros(0, np.int32)]).astype(np.int32)
    if ids.size == 0:
        ids = np.zeros(1, np.int32)
    return ptr, ids


def build_sparse_pbwt(P):
    """split_reference's PBWT build (product implementation)."""
    keep = _desc_arrays(P)
    d = _desc(P, keep)
    yl, al, nhq = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
    _check(lib().gh_build_sparse_pbwt(ctypes.byref(d), None, ctypes.byref(yl), None, None, ctypes.byref(al),
                                      ctypes.byref(nhq)))
    y = np.zeros(max(yl.value, 1), np.uint8)
    ap = np.zeros(nhq.value + 2, np.int32)
    ai = np.zeros(max(al.value, 1), np.int32)
    _check(lib().gh_build_sparse_pbwt(ctypes.byref(d), _p(y), ctypes.byref(yl), _p(ap), _p(ai), ctypes.byref(al),
                                      ctypes.byref(nhq)))
    return y[:yl.value], ap, ai[:al.value]


def _desc_arrays(P):
    return dict(cm=np.ascontiguousarray(P["cm"], np.float64), cref=np.ascontiguousarray(P["cref"], np.uint32),
                calt=np.ascontiguousarray(P["calt"], np.uint32),
                lq=np.ascontiguousarray(P.get("lq", np.zeros(P["L"], np.uint8)), np.uint8),
                cw=np.ascontiguousarray(P["common_words"], np.uint64),
                rp=np.ascontiguousarray(P["rare_ptr"], np.int32), ri=np.ascontiguousarray(P["rare_idx"], np.int32))


def _desc(P, k, ypacked=None, asmall=None):
    d = PanelDesc()
    d.n_sites, d.n_haps = P["L"], P["N"]
    d.cm, d.cref, d.calt, d.lq = _p(k["cm"]).value, _p(k["cref"]).value, _p(k["calt"]).value, _p(k["lq"]).value
    d.common_words, d.rare_ptr, d.rare_idx = _p(k["cw"]).value, _p(k["rp"]).value, _p(k["ri"]).value
    if ypacked is not None:
        d.ypacked, d.ypacked_len = _p(ypacked).value, ypacked.size
        d.asmall_ptr, d.asmall_idx = _p(asmall[0]).value, _p(asmall[1]).value
    d.sparse_maf = P.get("sparse_maf", 0.001)
    return d


class Panel:
    """A chunk panel resident on the GPU."""

    def __init__(self, P, with_pbwt=True):
        self.L, self.N = P["L"], P["N"]
        self._k = _desc_arrays(P)
        y = a = None
        if with_pbwt and "ypacked" in P:   # a split_reference panel carries its sparse-PBWT stream
            y = np.ascontiguousarray(P["ypacked"], np.uint8)
            ap = np.ascontiguousarray(P["asmall_ptr"], np.int32)
            ai = np.ascontiguousarray(P["asmall_idx"], np.int32)
            self._k.update(y=y, ap=ap, ai=ai)
            a = (ap, ai)
        elif with_pbwt:
            y, ap, ai = build_sparse_pbwt(P)
            self._k.update(y=y, ap=ap, ai=ai)
            a = (ap, ai)
        d = _desc(P, self._k, y, a)
        h = vp()
        _check(lib().gh_panel_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().gh_panel_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        nc, nh, nb = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        _check(lib().gh_panel_info(self.h, ctypes.byref(nc), ctypes.byref(nh), ctypes.byref(nb)))
        return dict(n_common=nc.value, n_common_hq=nh.value, device_bytes=nb.value)

    def condset_batch(self, state_lists, ne=100000):
        ptr, ids = _csr(state_lists)
        n = len(state_lists)
        cap_p = n * self.L
        poly_ptr = np.zeros(n + 1, np.int32)
        poly = np.zeros(cap_p, np.int32)
        hv_ptr = np.zeros(n + 1, np.int64)
        cap_w = sum(self.L * ((len(s) + 63) // 64) for s in state_lists)
        hv = np.zeros(cap_w, np.uint64)
        t = np.zeros(cap_p, np.float32)
        _check(lib().gh_condset_batch(self.h, n, _p(ptr), _p(ids), ne, _p(poly_ptr), _p(poly), cap_p, _p(hv_ptr),
                                      _p(hv), cap_w, _p(t)))
        out = []
        for i in range(n):
            W = (len(state_lists[i]) + 63) // 64
            p0, p1 = poly_ptr[i], poly_ptr[i + 1]
            Lp = p1 - p0
            out.append(dict(poly=poly[p0:p1].copy(), hvar=hv[hv_ptr[i]:hv_ptr[i + 1]].reshape(Lp, W).copy(),
                            t=t[p0 - i:p1 - i - 1].copy()))
        return out

    def init_states(self, gl, flat, kinit, seed, kstride=None):
        """gh_init_states: the INIT iteration's state sets of T diploid targets
        (gl [T][L][3], flat [T][L]); a list of ascending id arrays."""
        gl = np.ascontiguousarray(gl, np.uint8)
        flat = np.ascontiguousarray(flat, np.uint8)
        T = flat.shape[0]
        ks = int(kstride or min(kinit, self.N))
        st = np.zeros((T, ks), np.int32)
        K = np.zeros(T, np.int32)
        _check(lib().gh_init_states(self.h, T, _p(gl), _p(flat), int(kinit), int(seed) & 0xFFFFFFFF, ks, _p(st),
                                    _p(K)))
        return [st[t, :K[t]] for t in range(T)]

    def impute_batch(self, state_lists, HL, flat, err_imp=1e-12, ne=100000):
        ptr, ids = _csr(state_lists)
        n = len(state_lists)
        HL = np.ascontiguousarray(HL, np.float32).reshape(n, 2 * self.L)
        flat = np.ascontiguousarray(flat, np.uint8).reshape(n, self.L)
        HP = np.zeros((n, 2 * self.L), np.float32)
        _check(lib().gh_impute_batch(self.h, n, _p(ptr), _p(ids), _p(HL), _p(flat), ctypes.c_float(err_imp), ne,
                                     _p(HP)))
        return HP

    def rephase_batch(self, state_lists, flat, h0, h1, seed, it, tgt, err_phase=1e-4, ne=100000):
        ptr, ids = _csr(state_lists)
        n = len(state_lists)
        a = np.ascontiguousarray(h0, np.uint8).reshape(n, self.L).copy()
        b = np.ascontiguousarray(h1, np.uint8).reshape(n, self.L).copy()
        tg = np.ascontiguousarray(tgt, np.int32)
        _check(lib().gh_rephase_batch(self.h, n, _p(ptr), _p(ids), _p(np.ascontiguousarray(flat, np.uint8)), _p(a),
                                      _p(b), ctypes.c_float(err_phase), ne, seed, it, _p(tg)))
        return a, b

    def select_features(self, H, gl, flat, pool, pool_site, K, depth=12):
        """Learned-selector features (gh_select_features): H [2T][L], gl [T][L][3],
        flat [T][L], pool [T][kstride] (insertion-order codes), pool_site
        [T][kstride] (listing sites), K [T] -> X [T][kstride][32]."""
        pool_site = np.ascontiguousarray(pool_site, np.int32)
        H = np.ascontiguousarray(H, np.uint8)
        gl = np.ascontiguousarray(gl, np.uint8)
        flat = np.ascontiguousarray(flat, np.uint8)
        pool = np.ascontiguousarray(pool, np.int32)
        K = np.ascontiguousarray(K, np.int32)
        T, ks = pool.shape
        X = np.zeros((T, ks, MLP_IN), np.float32)
        _check(lib().gh_select_features(self.h, T, _p(H), _p(gl), _p(flat), _p(pool), _p(pool_site), _p(K), ks,
                                        int(depth), _p(X)))
        return X

    def pbwt_select(self, hap_alleles, stored, start_idx, depth=12, modulo=0.1, cap=None, kpbwt=0):
        """Returns the pbwt_states lists per target and, when kpbwt > 0, the
        compacted sorted conditioning set per target."""
        haps = np.ascontiguousarray(hap_alleles, np.uint8)
        T = haps.shape[0] // 2
        cap = cap or max(1 << 16, 64 * depth * T * 64)
        off = np.zeros(T * depth, np.int32)
        cnt = np.zeros(T * depth, np.int32)
        ids = np.zeros(cap, np.int32)
        st = np.zeros((T, max(kpbwt, 1)), np.int32) if kpbwt > 0 else None
        K = np.zeros(T, np.int32) if kpbwt > 0 else None
        _check(lib().gh_pbwt_select(self.h, T, _p(haps), _p(np.ascontiguousarray(stored, np.uint8)),
                                    _p(np.ascontiguousarray(start_idx, np.int32)), depth, ctypes.c_float(modulo),
                                    _p(off), _p(cnt), _p(ids), cap, int(kpbwt), _p(st), _p(K)))
        lists = [[ids[off[t * depth + d]: off[t * depth + d] + cnt[t * depth + d]].copy() for d in range(depth)]
                 for t in range(T)]
        if kpbwt > 0:
            return lists, [st[t, :K[t]].copy() for t in range(T)]
        return lists

    def phase_run(self, gl, flat, params, kpbwt_per_target=None):
        gl = np.ascontiguousarray(gl, np.uint8)
        flat = np.ascontiguousarray(flat, np.uint8)
        T, L = flat.shape
        n_iter = 1 + params.burnin + params.main_iter
        out = dict(gp0=np.zeros((T, L), np.float32), gp1=np.zeros((T, L), np.float32),
                   hds=np.zeros((T, L), np.uint8), h0=np.zeros((T, L), np.uint8), h1=np.zeros((T, L), np.uint8),
                   nstates=np.zeros((T, n_iter), np.int32))
        kpt = None if kpbwt_per_target is None else np.ascontiguousarray(kpbwt_per_target, np.int32)
        st = RunStats()
        _check(lib().gh_phase_run(self.h, T, _p(gl), _p(flat), ctypes.byref(params), _p(kpt), _p(out["gp0"]),
                                  _p(out["gp1"]), _p(out["hds"]), _p(out["h0"]), _p(out["h1"]),
                                  _p(out["nstates"]), ctypes.byref(st)))
        out["stats"] = {n: getattr(st, n) for n, _ in RunStats._fields_}
        return out


def output_fields(gp0, gp1, hds, cref, calt, ploidy=None):
    """gh_output_fields (gh_output_fields_ploidy with ploidy [T]): the writer's
    DS, GP, GT and per-site INFO / AF / RAF."""
    gp0 = np.ascontiguousarray(gp0, np.float32)
    gp1 = np.ascontiguousarray(gp1, np.float32)
    hds = np.ascontiguousarray(hds, np.uint8)
    cref = np.ascontiguousarray(cref, np.uint32)
    calt = np.ascontiguousarray(calt, np.uint32)
    T, L = gp0.shape
    out = dict(ds=np.zeros((T, L), np.float32), gp=np.zeros((T, L, 3), np.float32), gt=np.zeros((T, L, 2), np.uint8),
               info=np.zeros(L, np.float32), af=np.zeros(L, np.float32), raf=np.zeros(L, np.float32))
    if ploidy is None:
        _check(lib().gh_output_fields(T, L, _p(gp0), _p(gp1), _p(hds), _p(cref), _p(calt), _p(out["ds"]),
                                      _p(out["gp"]), _p(out["gt"]), _p(out["info"]), _p(out["af"]), _p(out["raf"])))
    else:
        pl = np.ascontiguousarray(np.broadcast_to(np.asarray(ploidy, np.int32), (T,)), np.int32)
        _check(lib().gh_output_fields_ploidy(T, L, _p(pl), _p(gp0), _p(gp1), _p(hds), _p(cref), _p(calt),
                                             _p(out["ds"]), _p(out["gp"]), _p(out["gt"]), _p(out["info"]),
                                             _p(out["af"]), _p(out["raf"])))
    return out


def select_topk(scores, pool, K, k_learn, states, min_mass=0.95):
    """gh_select_topk: returns (states [T][kstride], K_out [T], coverage [T])."""
    scores = np.ascontiguousarray(scores, np.float32)
    pool = np.ascontiguousarray(pool, np.int32)
    T, ks = pool.shape
    K = np.ascontiguousarray(K, np.int32)
    kl = np.ascontiguousarray(np.broadcast_to(np.asarray(k_learn, np.int32), (T,)), np.int32)
    st = np.array(states, np.int32, copy=True, order="C")
    Ko = np.zeros(T, np.int32)
    cov = np.zeros(T, np.float32)
    _check(lib().gh_select_topk(T, _p(scores), _p(pool), _p(K), _p(kl), ks, float(min_mass), _p(st), _p(Ko),
                                _p(cov)))
    return st, Ko, cov


class Workspace:
    """Device workspace shared by runs that execute one after another (gh_workspace_*)."""

    def __init__(self):
        h = vp()
        _check(lib().gh_workspace_create(ctypes.byref(h)))
        self.h = h

    @property
    def bytes(self):
        n = ctypes.c_int64()
        _check(lib().gh_workspace_bytes(self.h, ctypes.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().gh_workspace_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Run:
    """A prepared Gibbs run (gh_run_*): inputs resident on the device.
    execute() runs the whole INIT + burn-in + main schedule on the GPU and
    releases the GIL while it runs, so runs on different panels (streams) can
    be driven concurrently from Python threads."""

    def __init__(self, panel, gl, flat, params, kpbwt_per_target=None, ploidy=None):
        self.panel = panel
        self.gl = np.ascontiguousarray(gl, np.uint8)
        self.flat = np.ascontiguousarray(flat, np.uint8)
        self.T, self.L = self.flat.shape
        self.params = params
        self.n_iter = 1 + params.burnin + params.main_iter
        kpt = None if kpbwt_per_target is None else np.ascontiguousarray(kpbwt_per_target, np.int32)
        self._kpt = kpt
        h = vp()
        _check(lib().gh_run_create(panel.h, self.T, _p(self.gl), _p(self.flat), ctypes.byref(params), _p(kpt),
                                   ctypes.byref(h)))
        self.h = h
        if ploidy is not None:
            pl = np.ascontiguousarray(np.broadcast_to(np.asarray(ploidy, np.int32), (self.T,)), np.int32)
            _check(lib().gh_run_set_ploidy(self.h, _p(pl)))

    def execute(self):
        _check(lib().gh_run_execute(self.h))

    def set_tuning(self, mode=1):
        """Launch shapes (gh_run_set_tuning): 0 or False = throughput defaults,
        1 or True = autotune over the first executions, 2 or "latency" = latency
        defaults for the run on a batch's critical path."""
        m = 2 if mode == "latency" else int(mode)
        _check(lib().gh_run_set_tuning(self.h, m))

    def set_graph(self, mode=1):
        """gh_run_set_graph: 1 (default) = iterations from 2 on replay a
        captured HIP graph; 2 = the same device-decided iterations enqueued
        without a capture; 0 = every iteration launched from the host."""
        _check(lib().gh_run_set_graph(self.h, int(mode)))

    def set_workspace(self, ws):
        """Execute in a shared Workspace (None: the run's own); runs sharing one
        must not execute at the same time."""
        _check(lib().gh_run_set_workspace(self.h, ws.h if ws is not None else None))
        self._ws = ws

    def collect(self, from_iter=1):
        """Record (features, label, weight) rows for every PBWT pool candidate at
        iterations >= from_iter (a negative value switches recording off)."""
        _check(lib().gh_run_collect(self.h, int(from_iter)))

    def collected(self):
        n = ctypes.c_int64(0)
        _check(lib().gh_run_collected(self.h, ctypes.byref(n), None, None, None))
        X = np.zeros((n.value, MLP_IN), np.float32)
        y = np.zeros(n.value, np.float32)
        w = np.zeros(n.value, np.float32)
        _check(lib().gh_run_collected(self.h, ctypes.byref(n), _p(X), _p(y), _p(w)))
        return X, y, w

    def set_selector_prior(self, weight):
        """gh_run_set_selector_prior: rank score = predicted log mass - weight x
        (PBWT insertion position / pool size)."""
        _check(lib().gh_run_set_selector_prior(self.h, float(weight)))

    def set_selector(self, mlp, k_learn, min_mass=0.95, mode=1):
        """Use the learned selector: target t keeps k_learn[t] states of its PBWT
        pool, or the whole pool when their predicted mass share < min_mass.
        mode 0: best scores of the pool; mode 1: best scores per PBWT group,
        groups served round robin."""
        if mlp is None:
            _check(lib().gh_run_set_selector(self.h, None, None, 0.0, 0))
            self._mlp = None
            return
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(k_learn, np.int32), (self.T,)), np.int32)
        _check(lib().gh_run_set_selector(self.h, mlp.h, _p(k), float(min_mass), int(mode)))
        self._mlp = mlp   # keep the network alive while the run may use it

    def fetch(self):
        T, L = self.T, self.L
        out = dict(gp0=np.zeros((T, L), np.float32), gp1=np.zeros((T, L), np.float32),
                   hds=np.zeros((T, L), np.uint8), h0=np.zeros((T, L), np.uint8), h1=np.zeros((T, L), np.uint8),
                   nstates=np.zeros((T, self.n_iter), np.int32))
        st = RunStats()
        _check(lib().gh_run_fetch(self.h, _p(out["gp0"]), _p(out["gp1"]), _p(out["hds"]), _p(out["h0"]),
                                  _p(out["h1"]), _p(out["nstates"]), ctypes.byref(st)))
        out["stats"] = {n: getattr(st, n) for n, _ in RunStats._fields_}
        return out

    def stats(self):
        st = RunStats()
        _check(lib().gh_run_fetch(self.h, None, None, None, None, None, None, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in RunStats._fields_}

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().gh_run_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """The learned selector's network (gh_mlp_*): 32 -> 128 -> 128 -> 1, BF16
    operands on MFMA with FP32 accumulation, FP32 master weights and Adam."""

    def __init__(self, seed=15052011):
        h = vp()
        _check(lib().gh_mlp_create(int(seed) & 0xFFFFFFFF, ctypes.byref(h)))
        self.h = h

    def params(self):
        w = np.zeros(MLP_NPARAM, np.float32)
        _check(lib().gh_mlp_get_params(self.h, _p(w)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.shape == (MLP_NPARAM,)
        _check(lib().gh_mlp_set_params(self.h, _p(w)))

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_IN)
        y = np.zeros(X.shape[0], np.float32)
        _check(lib().gh_mlp_forward(self.h, X.shape[0], _p(X), _p(y)))
        return y

    def backward(self, X, t, w=None):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_IN)
        t = np.ascontiguousarray(t, np.float32)
        w = None if w is None else np.ascontiguousarray(w, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = ctypes.c_float(0)
        _check(lib().gh_mlp_backward(self.h, X.shape[0], _p(X), _p(t), _p(w), _p(g), ctypes.byref(loss)))
        return g, loss.value

    def step(self, X, t, w=None, lr=1e-3):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_IN)
        t = np.ascontiguousarray(t, np.float32)
        w = None if w is None else np.ascontiguousarray(w, np.float32)
        loss = ctypes.c_float(0)
        _check(lib().gh_mlp_step(self.h, X.shape[0], _p(X), _p(t), _p(w), float(lr), ctypes.byref(loss)))
        return loss.value

    def train_run(self, run, epochs=4, batch=16384, lr=2e-3, seed=1):
        """Adam on the rows a run recorded (device resident); returns the mean
        batch loss per epoch."""
        le = np.zeros(epochs, np.float32)
        _check(lib().gh_mlp_train_run(self.h, run.h, int(epochs), int(batch), float(lr), int(seed) & 0xFFFFFFFF,
                                      _p(le)))
        return le

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().gh_mlp_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
