"""Host side of the batched ERI path: device buffers (torch), plans, launches.

``EriPlan(batch)`` bins a ``basis.QuartetBatch`` by canonical class inside
libintc (C++), uploads the primitive data and S matrices, and ``run()`` executes
the pair builder and the class kernels on the current HIP stream.  Everything
numeric happens in libintc's kernels; this module only moves memory.
"""
import ctypes

import numpy as np

from . import lib as _lib
from .basis import SPH, ncc, nsph


def _torch():
    import torch
    return torch


def sph_blob(lmax=7):
    """S_l for l = 0..lmax concatenated column-major (the device layout)."""
    parts = []
    for l in range(lmax + 1):
        S = SPH[l] if l in SPH else None
        if S is None:
            from .basis import sph_matrix
            S = sph_matrix(l)
        parts.append(np.asfortranarray(S).reshape(-1, order="F"))
    return np.concatenate(parts)


def _cuda_device(torch, device):
    """torch.device with an explicit index ("cuda" -> the current device),
    which initialises torch's CUDA state before any device guard uses it."""
    d = torch.device(device)
    if d.type == "cuda" and d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class EriPlan:
    def __init__(self, batch, device="cuda", sph=None):
        """sph: None runs spherical plans with the built-in real solid
        harmonics (compiled into the kernels); a blob in the sph_blob()
        layout is uploaded and read by the kernels instead."""
        torch = _torch()
        L = _lib.load()
        self.batch = batch
        self.device = _cuda_device(torch, device)
        self._L = L
        b = batch
        sl = np.ascontiguousarray(b.shell_l, dtype=np.int32)
        snp = np.ascontiguousarray(b.shell_nprim, dtype=np.int32)
        soff = np.ascontiguousarray(b.shell_off, dtype=np.int32)
        cen = np.ascontiguousarray(b.shell_cen, dtype=np.float64)
        q = np.ascontiguousarray(b.quartets.reshape(-1), dtype=np.int32)
        oo = np.ascontiguousarray(b.out_off, dtype=np.int64)
        h = ctypes.c_void_p()
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        llp = oo.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
        ex = np.ascontiguousarray(b.exps, dtype=np.float64)
        co = np.ascontiguousarray(b.coefs, dtype=np.float64)
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        with torch.cuda.device(self.device):
            if b.pair_shells is not None:
                ps = np.ascontiguousarray(b.pair_shells.reshape(-1), dtype=np.int32)
                qp = np.ascontiguousarray(b.quartet_pairs.reshape(-1), dtype=np.int32)
                _lib.check(L.intc_plan_create_pairs(ctypes.byref(h), int(b.spherical), len(sl), ip(sl), ip(snp),
                                                    ip(soff), dp(cen), dp(ex), dp(co), len(ps) // 2, ip(ps),
                                                    len(oo), ip(qp), llp))
            else:
                _lib.check(L.intc_plan_create(ctypes.byref(h), int(b.spherical), len(sl), ip(sl), ip(snp),
                                              ip(soff), dp(cen), dp(ex), dp(co), len(oo), ip(q), llp))
        self.handle = h
        self.d_exps = torch.from_numpy(np.ascontiguousarray(b.exps, dtype=np.float64)).to(self.device)
        self.d_coefs = torch.from_numpy(np.ascontiguousarray(b.coefs, dtype=np.float64)).to(self.device)
        self.d_cen = torch.from_numpy(cen.reshape(-1)).to(self.device)
        self.d_sph = None
        if b.spherical and sph is not None:
            self.d_sph = torch.from_numpy(np.ascontiguousarray(sph, dtype=np.float64)).to(self.device)
        self.out_size = int(b.out_size)

    def bins(self, prec=0):
        """Class bins; prec 0 = FP64 kernels, 1 = FP32-recurrence kernels."""
        L = self._L
        n = L.intc_plan_num_bins(self.handle)
        out = []
        for i in range(n):
            cls = (ctypes.c_int * 4)()
            nq, T, grid, G = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            _lib.check(L.intc_plan_bin_info(self.handle, i, int(prec), cls, ctypes.byref(nq), ctypes.byref(T),
                                            ctypes.byref(grid), ctypes.byref(G)))
            out.append({"bin": i, "cls": tuple(cls), "nq": nq.value, "team_groups": T.value,
                        "grid": grid.value, "G": G.value, "prec": prec})
        return out

    def set_tuning(self, i, team_groups, grid, prec=0):
        _lib.check(self._L.intc_plan_set_bin_tuning(self.handle, int(i), int(prec), int(team_groups), int(grid)))

    def _stream(self):
        torch = _torch()
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def alloc_out(self):
        torch = _torch()
        return torch.empty(self.out_size, dtype=torch.float64, device=self.device)

    def build_pairs(self):
        _lib.check(self._L.intc_plan_build_pairs(self.handle, _ptr(self.d_exps), _ptr(self.d_coefs),
                                                 _ptr(self.d_cen), self._stream()))

    def run_bin(self, i, out, prec=0):
        sph = _ptr(self.d_sph) if self.d_sph is not None else ctypes.c_void_p()
        _lib.check(self._L.intc_plan_run_bin(self.handle, int(i), int(prec), sph, _ptr(out), self._stream()))

    def run(self, out=None):
        if out is None:
            out = self.alloc_out()
        sph = _ptr(self.d_sph) if self.d_sph is not None else ctypes.c_void_p()
        _lib.check(self._L.intc_plan_execute(self.handle, _ptr(self.d_exps), _ptr(self.d_coefs),
                                             _ptr(self.d_cen), sph, _ptr(out), self._stream()))
        return out

    def run_routed(self, route, out=None):
        """route: device int8 tensor, one code per quartet (0 FP64, 1 FP32, 2 skip)."""
        if out is None:
            out = self.alloc_out()
        sph = _ptr(self.d_sph) if self.d_sph is not None else ctypes.c_void_p()
        _lib.check(self._L.intc_plan_execute_routed(self.handle, _ptr(self.d_exps), _ptr(self.d_coefs),
                                                    _ptr(self.d_cen), sph, _ptr(route), _ptr(out), self._stream()))
        return out

    def set_route_fp32(self, use_fp32):
        _lib.check(self._L.intc_plan_set_route_fp32(self.handle, int(bool(use_fp32))))

    def set_fp32_check(self, tol):
        """Routed runs: the FP32-recurrence register kernels estimate each
        quartet's error (8 * 2^(la+lb+lc+ld) * 2^-24 times its largest
        contracted intermediate) and queue the quartets above tol for the
        FP64 kernel, which recomputes them after the FP32 launch.  0: off."""
        _lib.check(self._L.intc_plan_set_fp32_check(self.handle, float(tol)))

    def redo_counts(self):
        """Per bin: quartets the last routed run recomputed in FP64."""
        n = self._L.intc_plan_num_bins(self.handle)
        buf = (ctypes.c_int * max(n, 1))()
        _lib.check(self._L.intc_plan_redo_counts(self.handle, buf))
        return np.array(buf[:n])

    def set_persistent(self, enable=True, tile_iters=64, grid=0):
        """One persistent launch for every bin whose class the multi-class
        kernel holds (spherical, built-in transforms, Chebyshev Boys): waves
        take tiles of (teams per wave) x (at most tile_iters) quartet
        positions from a device counter, bins interleaved.  grid = 0: two
        blocks of 256 per CU.  Applies to run, run_routed and later
        captures."""
        _lib.check(self._L.intc_plan_set_persistent(self.handle, int(bool(enable)), int(tile_iters), int(grid)))

    def set_persistent_bins(self, cover=None):
        """Bins the persistent launch may hold: a boolean per bin (bins()
        order), or None for every bin its kernel holds.  The others keep their
        own class kernels."""
        if cover is None:
            _lib.check(self._L.intc_plan_set_persistent_bins(self.handle, 0, None))
            return
        c = np.ascontiguousarray(np.asarray(cover, dtype=bool).astype(np.int32))
        _lib.check(self._L.intc_plan_set_persistent_bins(self.handle, len(c),
                                                         c.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def persistent_info(self):
        """(tiles, LDS doubles per wave) of the last persistent eager run."""
        n, r = ctypes.c_int(), ctypes.c_int()
        _lib.check(self._L.intc_plan_persistent_info(self.handle, ctypes.byref(n), ctypes.byref(r)))
        return n.value, r.value

    def set_prim_screen(self, cut):
        """Keep only primitive pairs with s-type Schwarz factor w >= cut in
        builds after this call (0 keeps every pair); see prim_screen_cut."""
        _lib.check(self._L.intc_plan_set_prim_screen(self.handle, float(cut)))

    def pair_prims(self, screened=False):
        """Primitive pairs per shell pair (plan pair order): all, or those
        the last screened build kept."""
        n = int(self._n_pair())
        out = np.zeros(max(n, 1), dtype=np.int32)
        _lib.check(self._L.intc_plan_pair_prims(self.handle, int(bool(screened)),
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out[:n]

    def _n_pair(self):
        # a plan over shared pairs keeps the caller's pairs; otherwise every
        # quartet has its own bra and ket pair
        b = self.batch
        return len(b.pair_shells) if b.pair_shells is not None else 2 * b.n_quartet

    def prim_screen_cut(self, eps):
        """Primitive cut eps / W_max, W_max the largest sum of w over one
        shell pair of this batch (pair_w_sums).

        What the cut bounds: dropping primitive pairs with w < cut moves a
        contracted integral by at most A (D_bra W_ket + W_bra D_ket), D a
        pair's dropped sum of w (at most n_dropped * cut) and A an angular
        factor (1 for s shells; at most 7 measured on 3000 config-3 quartets
        with l <= 2, no closed form).  With this cut the change is therefore
        up to A (n_bra + n_ket) eps <= 2 A Ka Kb eps, not eps: 32 A eps for
        K = 4 shells.  Callers with a hard error budget measure the result
        (the benchmark does) or pass eps / (2 Ka Kb A)."""
        return float(eps) / max(pair_w_sums(self.batch).max(initial=0.0), 1e-300)

    def route_counts(self):
        n = self._L.intc_plan_num_bins(self.handle)
        buf = (ctypes.c_int * (3 * max(n, 1)))()
        _lib.check(self._L.intc_plan_route_counts(self.handle, buf))
        return np.array(buf[:3 * n]).reshape(n, 3)

    def quartet_order(self):
        """order[i] = index (in the batch) of the quartet at plan position i."""
        order = np.empty(self.batch.n_quartet, dtype=np.int32)
        _lib.check(self._L.intc_plan_quartet_order(self.handle, order.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return order

    def capture(self, out, n_streams=4, route=None):
        """Capture pair building (+ route compaction) + all bins into one
        hipGraph bound to `out` (and `route`)."""
        sph = _ptr(self.d_sph) if self.d_sph is not None else ctypes.c_void_p()
        r = _ptr(route) if route is not None else ctypes.c_void_p()
        _lib.check(self._L.intc_plan_graph_create(self.handle, _ptr(self.d_exps), _ptr(self.d_coefs),
                                                  _ptr(self.d_cen), sph, r, _ptr(out), int(n_streams),
                                                  self._stream()))
        self._graph_out = out
        return out

    def replay(self):
        _lib.check(self._L.intc_plan_graph_launch(self.handle, self._stream()))
        return self._graph_out

    def close(self):
        if self.handle:
            self._L.intc_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GeneralEriPlan(EriPlan):
    """A plan over general contractions with the caller's strides (the
    per-quartet fields of the reference's <full>_main for a batch).  shells:
    list of dicts with l, exps (n_prim), coefs (n_prim x n_cont, normalised)
    and cen; quartets (n, 4) shell indices; out_off[q], iskip[q] (4) and
    iskip_cont[q] (4) place element (ia, ib, ic, id) of contraction
    (ka, kb, kc, kd) at out[out_off[q] + sum_X(iX iskip_X + kX iskip_cont_X)].
    out_size: length of the output array.  run(), build_pairs(), run_bin()
    and capture() as for EriPlan; routed execution is refused."""

    def __init__(self, spherical, shells, quartets, out_off, iskip, iskip_cont, out_size, device="cuda", sph=None):
        torch = _torch()
        L = _lib.load()
        self._L = L
        self.batch = None
        self.device = _cuda_device(torch, device)
        ls = np.array([s["l"] for s in shells], dtype=np.int32)
        cols = [np.asarray(s["coefs"], dtype=np.float64).reshape(len(s["exps"]), -1) for s in shells]
        nprim = np.array([len(s["exps"]) for s in shells], dtype=np.int32)
        ncont = np.array([c.shape[1] for c in cols], dtype=np.int32)
        soff = np.concatenate([[0], np.cumsum(nprim)[:-1]]).astype(np.int32)
        coff = np.concatenate([[0], np.cumsum(nprim * ncont)[:-1]]).astype(np.int32)
        ex = np.ascontiguousarray(np.concatenate([np.asarray(s["exps"], dtype=np.float64) for s in shells]))
        co = np.ascontiguousarray(np.concatenate([c.reshape(-1, order="F") for c in cols]))
        cen = np.ascontiguousarray(np.array([s["cen"] for s in shells], dtype=np.float64).reshape(-1))
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.int32).reshape(-1))
        oo = np.ascontiguousarray(np.asarray(out_off, dtype=np.int64))
        isk = np.ascontiguousarray(np.asarray(iskip, dtype=np.int32).reshape(-1))
        iskc = np.ascontiguousarray(np.asarray(iskip_cont, dtype=np.int32).reshape(-1))
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(L.intc_plan_create_general(ctypes.byref(h), int(bool(spherical)), len(ls), ip(ls), ip(nprim),
                                                  ip(ncont), ip(soff), ip(coff), dp(cen), dp(ex), dp(co), len(oo),
                                                  ip(q), oo.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                                  ip(isk), ip(iskc)))
        self.handle = h
        self.d_exps = torch.from_numpy(ex).to(self.device)
        self.d_coefs = torch.from_numpy(co).to(self.device)
        self.d_cen = torch.from_numpy(cen).to(self.device)
        self.d_sph = None
        if spherical and sph is not None:
            self.d_sph = torch.from_numpy(np.ascontiguousarray(sph, dtype=np.float64)).to(self.device)
        self.out_size = int(out_size)

    def run_routed(self, route, out=None):
        raise NotImplementedError("routed execution needs dense output blocks")


class OneEPlan:
    """Batched one-electron integrals: kind 0 overlap (items = shell pairs
    (a, b)), kind 1 nuclear attraction (items = (a, b, c), c indexing
    `centres`, the point charges).  shells: any object with shell_l,
    shell_nprim, shell_off, shell_cen, exps and coefs (e.g. a QuartetBatch or
    a Molecule); blocks are row-major [a][b], dense in item order unless
    out_off is given.  sph: None = built-in real solid harmonics."""

    def __init__(self, kind, shells, items, centres=None, spherical=True, out_off=None, device="cuda", sph=None):
        torch = _torch()
        L = _lib.load()
        self._L = L
        self.kind = int(kind)
        self.device = _cuda_device(torch, device)
        items = np.ascontiguousarray(items, dtype=np.int32)
        w = 2 if self.kind == 0 else 3
        assert items.ndim == 2 and items.shape[1] == w
        sl = np.ascontiguousarray(shells.shell_l, dtype=np.int32)
        snp = np.ascontiguousarray(shells.shell_nprim, dtype=np.int32)
        soff = np.ascontiguousarray(shells.shell_off, dtype=np.int32)
        nc = np.array([(2 * l + 1) if spherical else (l + 1) * (l + 2) // 2 for l in range(8)], dtype=np.int64)
        self.sizes = nc[sl[items[:, 0]]] * nc[sl[items[:, 1]]]
        if out_off is None:
            out_off = np.zeros(len(items), dtype=np.int64)
            if len(items):
                out_off[1:] = np.cumsum(self.sizes)[:-1]
        self.out_off = np.ascontiguousarray(out_off, dtype=np.int64)
        self.out_size = int((self.out_off + self.sizes).max()) if len(items) else 0
        n_centre = 0 if centres is None else len(centres)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        h_cen = np.ascontiguousarray(shells.shell_cen, dtype=np.float64).reshape(-1)
        h_ex = np.ascontiguousarray(shells.exps, dtype=np.float64)
        h_co = np.ascontiguousarray(shells.coefs, dtype=np.float64)
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(L.intc_onee_plan_create(ctypes.byref(h), self.kind, int(bool(spherical)), len(sl), ip(sl),
                                               ip(snp), ip(soff), dp(h_cen), dp(h_ex), dp(h_co), n_centre,
                                               len(items), ip(items),
                                               self.out_off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
        self.handle = h
        f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).reshape(-1)).to(self.device)
        self.d_exps, self.d_coefs, self.d_cen = f64(shells.exps), f64(shells.coefs), f64(shells.shell_cen)
        self.d_centres = f64(centres) if centres is not None else None
        self.d_sph = f64(sph) if sph is not None else None

    def run(self, out=None):
        torch = _torch()
        if out is None:
            out = torch.empty(self.out_size, dtype=torch.float64, device=self.device)
        c = _ptr(self.d_centres) if self.d_centres is not None else ctypes.c_void_p()
        sph = _ptr(self.d_sph) if self.d_sph is not None else ctypes.c_void_p()
        st = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(self._L.intc_onee_plan_execute(self.handle, _ptr(self.d_exps), _ptr(self.d_coefs),
                                                  _ptr(self.d_cen), c, sph, _ptr(out), st))
        return out

    def close(self):
        if self.handle:
            self._L.intc_onee_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pair_w_sums(batch):
    """Sum over primitive pairs of the s-type Schwarz factor
    w = |c_a c_b| K_AB sqrt(2 pi^(5/2)) / zeta (2 zeta)^(-1/4) for every shell
    pair of the batch (its shared pairs, else the bra and ket of each quartet),
    the factor the device primitive screen compares with its cut."""
    b = batch
    if b.pair_shells is not None:
        pairs = np.asarray(b.pair_shells, dtype=np.int64).reshape(-1, 2)
    else:
        q = np.asarray(b.quartets, dtype=np.int64).reshape(-1, 4)
        pairs = np.unique(np.sort(np.concatenate([q[:, :2], q[:, 2:]]), axis=1), axis=0)
    sa, sb = pairs[:, 0], pairs[:, 1]
    cen = np.asarray(b.shell_cen, dtype=np.float64).reshape(-1, 3)
    r2 = np.sum((cen[sa] - cen[sb]) ** 2, axis=1)
    na, nb = np.asarray(b.shell_nprim)[sa], np.asarray(b.shell_nprim)[sb]
    oa, ob = np.asarray(b.shell_off)[sa], np.asarray(b.shell_off)[sb]
    tot = np.zeros(len(pairs))
    for i in range(int(na.max(initial=0))):
        for j in range(int(nb.max(initial=0))):
            m = (i < na) & (j < nb)
            al, be = b.exps[oa[m] + i], b.exps[ob[m] + j]
            z = al + be
            kap = 5.914967172795612878 * np.exp(-al * be / z * r2[m]) * b.coefs[oa[m] + i] * b.coefs[ob[m] + j] / z
            tot[m] += np.abs(kap) * (2.0 * z) ** -0.25
    return tot


def set_boys_mode(mode):
    _lib.check(_lib.load().intc_set_boys_mode(int(mode)))


def compute(batch, device="cuda", sph=None):
    """Convenience: all integrals of a batch as a host numpy array (sph as
    for EriPlan)."""
    torch = _torch()
    plan = EriPlan(batch, device, sph=sph)
    out = plan.run()
    torch.cuda.synchronize(plan.device)
    res = out.cpu().numpy()
    plan.close()
    return res
