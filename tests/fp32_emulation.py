"""TEST INFRASTRUCTURE ONLY: a numpy restatement of the register class
kernels' recurrence arithmetic, for the FP32 error model
(tests/test_fp32_error_model.py).

The recurrence lines are the kernel generator's own (codegen._chain_lines for
the bra chain, codegen._ket_lines for the ket recursion), translated to numpy
and evaluated over every primitive quartet at once; axis 0 is the bra
component (the kernels' lane), and the cross-lane term reads the row of
component e - 1_i.  Precision "f" follows the FP32 kernels: FP64 setup (Boys
values, prefactors, coefficients), FP32 recurrences on s^m-scaled values
(s = max(1, 2T)), FP64 contraction sums; "d" is all FP64.  The contracted
(e 0|f 0) then go through an FP64 HGP HRR and the built-in solid-harmonic
transforms.  numpy rounds every product and sum separately (no FMA
contraction), so the errors are of the kernels' size, not bit-identical.

mode (FP32 only):
  None            plain FP32 recurrences;
  "step_rounded"  every recurrence value computed exactly from its FP32
                  inputs and rounded to FP32 once: the limit of any
                  compensated (error-free transformation) evaluation of the
                  sums of one recurrence step;
  "inputs_only"   exact recurrences on FP32-rounded inputs (coefficients and
                  scaled Boys values): the part of the error that no
                  arithmetic on FP32 inputs can remove.
"""
import re

import numpy as np

from intception_amd import codegen as cg
from intception_amd.basis import SPH, cart_components, ncc, nsph, sph_matrix
from oracle import oracle

_SEL = re.compile(r"(\w+) == 0 \? (\w+) : \(\1 == 1 \? (\w+) : \(\1 == 2 \? (\w+) : ([\w.]+)\)\)")


def _translate(line):
    s = line.strip().rstrip(";")
    s = re.sub(r"^const (int|float|double) ", "", s)
    s = _SEL.sub(lambda m: "SEL(%s, %s, %s, %s, %s)" % m.groups(), s)
    s = re.sub(r"\((float|double)\)\(", r"CAST(", s)
    s = re.sub(r"(\d+\.\d*)f\b", r"F(\1)", s)
    s = re.sub(r"(?<![\w.])(\d+\.\d*)(?![\w.])", r"F(\1)", s)
    return s.replace("intc_shflf(", "SHFL(").replace("intc_shfl(", "SHFL(")


def _lane_tab(LX):
    """components, cross-lane sources, chain directions and kk multipliers,
    with the kernels' padding at the front of the chain (the value before the
    last is the lane's predecessor, which codegen._ket_lines reads for the
    first ket level)"""
    rows = []
    for e in cg.comps_upto(LX):
        src = [cg.cidx(cg.dec(e, i)) if e[i] > 0 else cg.cidx(e) for i in range(3)]
        steps, t = [], e
        while sum(t) > 0:
            i = cg.first_nonzero(t)
            steps.append((i, t[i] - 1))
            t = cg.dec(t, i)
        steps = steps[::-1]
        pad = LX - len(steps)
        dirs = ks = 0
        for s in range(LX):
            d, k = (3, 0) if s < pad else steps[s - pad]
            dirs |= d << (2 * s)
            ks |= k << (4 * s)
        rows.append((e[0], e[1], e[2], src[0], src[1], src[2], dirs, ks))
    return np.array(rows, dtype=np.int64)


def setup(b, cls):
    """FP64 primitive-quartet parameters of a batch whose quartets are in the
    canonical orientation of cls"""
    q = b.quartets
    rows = []
    for qi, (sa, sb, sc, sd) in enumerate(q):
        o = [b.shell_off[s] for s in (sa, sb, sc, sd)]
        n = [b.shell_nprim[s] for s in (sa, sb, sc, sd)]
        for ia in range(n[0]):
            for ib in range(n[1]):
                for ic in range(n[2]):
                    for id_ in range(n[3]):
                        rows.append((qi, o[0] + ia, o[1] + ib, o[2] + ic, o[3] + id_))
    r = np.array(rows, dtype=np.int64)
    qi = r[:, 0]
    ex, co = b.exps, b.coefs
    A, B = b.shell_cen[q[qi, 0]], b.shell_cen[q[qi, 1]]
    C, D = b.shell_cen[q[qi, 2]], b.shell_cen[q[qi, 3]]
    al, be, ga, de = ex[r[:, 1]], ex[r[:, 2]], ex[r[:, 3]], ex[r[:, 4]]
    zeta, eta = al + be, ga + de
    P = (al[:, None] * A + be[:, None] * B) / zeta[:, None]
    Q = (ga[:, None] * C + de[:, None] * D) / eta[:, None]
    K = np.sqrt(2.0 * np.pi ** 2.5)
    kab = K * np.exp(-al * be / zeta * np.sum((A - B) ** 2, 1)) * co[r[:, 1]] * co[r[:, 2]] / zeta
    kcd = K * np.exp(-ga * de / eta * np.sum((C - D) ** 2, 1)) * co[r[:, 3]] * co[r[:, 4]] / eta
    return dict(qi=qi, nq=len(q), A=A, C=C, P=P, Q=Q, zeta=zeta, eta=eta, kab=kab, kcd=kcd)


def contracted(S, cls, prec, mode=None):
    """{f: (G, nq)} contracted (e 0|f 0) for |f| >= lc (rows: every bra
    component), from the kernels' recurrence lines"""
    LA, LB, LC, LD = cls
    LX, LR = LA + LB, LC + LD
    M = LX + LR
    R = np.float32 if prec == "f" else np.float64
    zeta, eta, P, Q = S["zeta"], S["eta"], S["P"], S["Q"]
    os_ = 1.0 / (zeta + eta)
    PQ = P - Q
    T = zeta * eta * os_ * np.sum(PQ * PQ, 1)
    pref = S["kab"] * S["kcd"] * np.sqrt(os_)
    F = np.stack([oracle.boys_exact(M, float(t)) for t in T], axis=1)
    sc = np.where(T > 0.5, 2.0 * T, 1.0) if prec == "f" else np.ones_like(T)
    isc = 1.0 / sc
    eo, zo = eta * os_, zeta * os_
    PA, QC = P - S["A"], Q - S["C"]
    exact = {}
    for i, ax in enumerate("xyz"):
        exact["PA" + ax], exact["WP" + ax] = PA[:, i], -eo * PQ[:, i] * isc
        exact["QC" + ax], exact["WQ" + ax] = QC[:, i], zo * PQ[:, i] * isc
    exact.update(roz=eo * isc, roe=zo * isc, o2s=0.5 * os_ * isc, o2z=0.5 / zeta, o2e=0.5 / eta)
    exact["o2zf"], exact["o2ef"] = exact["o2z"], exact["o2e"]
    scm = np.ones_like(T)
    for m in range(M + 1):
        exact["c0_%d" % m] = F[m] * scm if prec == "f" else pref * F[m]
        scm = scm * sc
    lt = _lane_tab(LX)
    G = len(lt)
    work = np.float64 if mode in ("step_rounded", "inputs_only") else R
    env = {k: np.asarray(v).astype(R).astype(work)[None, :] for k, v in exact.items()}
    env["cdir"], env["ckk"] = lt[:, 6][:, None], lt[:, 7][:, None]
    env["ex_d"], env["ey_d"], env["ez_d"] = (lt[:, k].astype(work)[:, None] for k in range(3))
    env["s4x"], env["s4y"], env["s4z"] = lt[:, 3], lt[:, 4], lt[:, 5]
    env["SEL"] = lambda d, x, y, z, c: np.where(d == 0, x, np.where(d == 1, y, np.where(d == 2, z, c)))
    env["SHFL"] = lambda v, src: v[src]
    env["F"] = work
    env["CAST"] = lambda x: np.asarray(x).astype(work)
    lines = []
    if LX > 0 and LR > 0:
        lines += ["cplx = ex_d * o2s", "cply = ey_d * o2s", "cplz = ez_d * o2s"]
    Rname = "float" if prec == "f" else "double"
    body = cg._chain_lines(LX, M, Rname, None)
    kl, V = cg._ket_lines(LX, LR, Rname)
    lines += [_translate(ln) for ln in body + kl]
    for ln in lines:
        name = ln.split("=")[0].strip()
        exec(ln, env)
        if mode == "step_rounded" and isinstance(env[name], np.ndarray) and env[name].dtype == np.float64:
            env[name] = env[name].astype(np.float32).astype(np.float64)
    out = {}
    for f in [f for f in cg.comps_upto(LR) if sum(f) >= LC]:
        v = np.broadcast_to(np.asarray(env[V(f, 0)], dtype=np.float64), (G, len(T)))
        contrib = v * (pref[None, :] if prec == "f" else 1.0)
        acc = np.zeros((G, S["nq"]))
        np.add.at(acc.T, S["qi"], contrib.T)
        out[f] = acc
    return out


def hrr_sph(out, b, cls):
    """FP64 HGP HRR (ket, then bra) and the built-in transforms: (nq, block)"""
    LA, LB, LC, LD = cls
    LX = LA + LB
    q = b.quartets
    AB = b.shell_cen[q[:, 0]] - b.shell_cen[q[:, 1]]
    CD = b.shell_cen[q[:, 2]] - b.shell_cen[q[:, 3]]

    def hrr(src, lo, nmove, Rv):
        vals = {(e, (0, 0, 0)): v for e, v in src.items()}
        for bl in range(1, nmove + 1):
            for e in cg.comps_upto(lo + nmove - bl):
                if sum(e) < lo:
                    continue
                for bb in cart_components(bl):
                    i = cg.first_nonzero(bb)
                    b1 = cg.dec(bb, i)
                    e1 = list(e)
                    e1[i] += 1
                    vals[(e, bb)] = vals[(tuple(e1), b1)] + Rv[:, i] * vals[(e, b1)]
        return vals
    Ss = [SPH[l] if l in SPH else sph_matrix(l) for l in cls]
    nq = len(q)
    res = np.zeros((nq, nsph(LA), nsph(LB), nsph(LC), nsph(LD)))
    ket = {}
    for ei, e in enumerate(cg.comps_upto(LX)):
        if sum(e) < LA:
            continue
        hv = hrr({f: out[f][ei] for f in out}, LC, LD, CD)
        X = np.zeros((nq, ncc(LC), ncc(LD)))
        for ci, c in enumerate(cart_components(LC)):
            for di, d in enumerate(cart_components(LD)):
                X[:, ci, di] = hv[(c, d)]
        ket[e] = np.einsum("qcd,cs,dt->qst", X, Ss[2], Ss[3])
    for s in range(nsph(LC)):
        for t in range(nsph(LD)):
            hv = hrr({e: ket[e][:, s, t] for e in ket}, LA, LB, AB)
            X = np.zeros((nq, ncc(LA), ncc(LB)))
            for ai, a in enumerate(cart_components(LA)):
                for bi, bb in enumerate(cart_components(LB)):
                    X[:, ai, bi] = hv[(a, bb)]
            res[:, :, :, s, t] = np.einsum("qab,as,bt->qst", X, Ss[0], Ss[1])
    return res.reshape(nq, -1)


def evaluate(b, cls, prec, mode=None):
    """(spherical blocks (nq, block), largest |contracted (e 0|f 0)| per quartet)"""
    S = setup(b, cls)
    out = contracted(S, cls, prec, mode)
    keep = [k for k, e in enumerate(cg.comps_upto(cls[0] + cls[1])) if sum(e) >= cls[0]]
    inter = np.max([np.abs(a[keep]).max(0) for a in out.values()], axis=0)
    return hrr_sph(out, b, cls), inter
