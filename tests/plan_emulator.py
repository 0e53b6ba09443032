"""CPU emulation of the device algorithm on a symbolic plan (kg_plan_export).

Test infrastructure: it runs the same gather / factor-program / level-order
solve steps the gfx950 kernels run, in numpy, so a plan can be checked on a
machine without a GPU.
"""
import ctypes
import math

import numpy as np

from krets_amd import _lib

FIELDS = ["n", "nnz", "ne", "nlev", "nblev", "ntlev", "ntblev", "nonlinear", "has_diode", "has_dyn",
          "o_row_ptr", "o_col", "o_diag", "o_fprog_ptr", "o_fprog", "o_lvl_ptr", "o_lvl_rows", "o_blvl_ptr",
          "o_blvl_rows", "o_scon_ptr", "o_scon", "o_rcon_ptr", "o_rcon", "o_col_of", "o_cpos", "o_isbr",
          "o_etype", "o_ep", "o_em", "o_ec", "o_ut_ptr", "o_ut", "o_lt_ptr", "o_lt", "o_tlvl_ptr", "o_tlvl_rows",
          "o_tblvl_ptr", "o_tblvl_rows", "o_row_of", "o_rpos", "o_erp", "o_erm", "o_erc", "status", "total"]
DC, AC, TRAN = 0, 1, 2
NATURAL, MINDEG, NESTED, BALANCED = 0, 1, 2, 3


class Plan:
    def __init__(self, elements, n, kind, ordering, params=None, s_rep=1.0):
        el = np.ascontiguousarray(elements, _lib.ELEMENT_DTYPE)
        lay = np.zeros(64, np.int32)
        L = _lib.lib()
        pr = None if params is None else np.ascontiguousarray(params, np.float64)
        pp = None if pr is None else pr.ctypes.data
        size = L.kg_plan_export(el.ctypes.data, len(el), n, kind, ordering, pp, s_rep, None, 0, lay.ctypes.data, 64)
        self.status = -size if size < 0 else 0
        if size < 0:
            return
        buf = np.zeros(size, np.int32)
        L.kg_plan_export(el.ctypes.data, len(el), n, kind, ordering, pp, s_rep, buf.ctypes.data, size,
                         lay.ctypes.data, 64)
        self.L = {k: int(lay[i]) for i, k in enumerate(FIELDS)}
        self.b = buf
        self.n, self.nnz, self.ne = self.L["n"], self.L["nnz"], self.L["ne"]

    def arr(self, name, length):
        o = self.L["o_" + name]
        return self.b[o:o + length]

    @property
    def row_ptr(self):
        return self.arr("row_ptr", self.n + 1)

    @property
    def col(self):
        return self.arr("col", self.nnz)

    @property
    def diag(self):
        return self.arr("diag", self.n)

    def levels(self, which):
        nl = self.L[{"lvl": "nlev", "blvl": "nblev", "tlvl": "ntlev", "tblvl": "ntblev"}[which]]
        ptr = self.arr(which + "_ptr", nl + 1)
        rows = self.arr(which + "_rows", self.n)
        return [list(rows[ptr[l]:ptr[l + 1]]) for l in range(nl)]

    def gather(self, contrib_value):
        """Dense permuted matrix from the per-slot contribution lists."""
        scp = self.arr("scon_ptr", self.nnz + 1)
        sc = self.arr("scon", int(scp[-1]))
        rp, col = self.row_ptr, self.col
        v = np.zeros(self.nnz, complex)
        for s in range(self.nnz):
            v[s] = sum(contrib_value(int(c) >> 3, int(c) & 7) for c in sc[scp[s]:scp[s + 1]])
        M = np.zeros((self.n, self.n), complex)
        for i in range(self.n):
            for p in range(rp[i], rp[i + 1]):
                M[i, col[p]] = v[p]
        return v, M

    def gather_rhs(self, rhs_value):
        rcp = self.arr("rcon_ptr", self.n + 1)
        rc = self.arr("rcon", int(rcp[-1]))
        b = np.zeros(self.n, complex)
        for i in range(self.n):
            b[i] = sum(rhs_value(int(c) >> 3, int(c) & 7) for c in rc[rcp[i]:rcp[i + 1]])
        return b

    def factor(self, v):
        """Runs the factor program level by level; rows inside a level in
        reverse order (any order must give the same result)."""
        v = v.copy()
        fpp = self.arr("fprog_ptr", self.n + 1)
        fp = self.b[self.L["o_fprog"]:]
        diag = self.diag
        for lev in self.levels("lvl"):
            for i in reversed(lev):
                q = fpp[i]
                nk = fp[q]
                q += 1
                for _ in range(nk):
                    sik, sdk, nu = fp[q], fp[q + 1], fp[q + 2]
                    q += 3
                    lik = v[sik] * v[sdk]
                    v[sik] = lik
                    for _ in range(nu):
                        v[fp[q]] -= lik * v[fp[q + 1]]
                        q += 2
                v[diag[i]] = 1.0 / v[diag[i]]
        return v

    def solve(self, v, b):
        w = b.astype(complex).copy()
        rp, col, diag = self.row_ptr, self.col, self.diag
        for lev in self.levels("lvl"):
            for i in reversed(lev):
                w[i] -= sum(v[p] * w[col[p]] for p in range(rp[i], diag[i]))
        for lev in self.levels("blvl"):
            for i in reversed(lev):
                w[i] = (w[i] - sum(v[p] * w[col[p]] for p in range(diag[i] + 1, rp[i + 1]))) * v[diag[i]]
        return w

    def solve_t(self, v, b):
        w = b.astype(complex).copy()
        utp = self.arr("ut_ptr", self.n + 1)
        ut = self.b[self.L["o_ut"]:]
        ltp = self.arr("lt_ptr", self.n + 1)
        lt = self.b[self.L["o_lt"]:]
        diag = self.diag
        for lev in self.levels("tlvl"):
            for i in reversed(lev):
                s = w[i] - sum(v[ut[q + 1]] * w[ut[q]] for q in range(utp[i], utp[i + 1], 2))
                w[i] = s * v[diag[i]]
        for lev in self.levels("tblvl"):
            for i in reversed(lev):
                w[i] -= sum(v[lt[q + 1]] * w[lt[q]] for q in range(ltp[i], ltp[i + 1], 2))
        return w


def guard_bound_coeff(m):
    """Backward-error bound per unit growth of an LU without pivoting over
    rows of at most m entries (kg_internal.h guard_bound_coeff)."""
    u = 2.0 ** -53
    g2 = 2 * u / (1 - 2 * u)
    q = m * np.sqrt(2) * g2
    return 3.01 * np.sqrt(2) * q / (1 - q)


def static_pivot_bound(plan, contrib):
    """The static-pivot guard's bound guard_bound_coeff(m) * rho,
    rho = || |L||U| ||_inf / ||A||_inf, for the plan's pivot order, from a
    dense elimination without pivoting of the plan's permuted matrix in numpy
    (m: the longest row of the factor pattern); entries |re| + |im|."""
    _, M = plan.gather(contrib)
    n = M.shape[0]
    L, U = np.eye(n, dtype=complex), M.copy()
    for j in range(n):
        L[j + 1:, j] = U[j + 1:, j] / U[j, j]
        U[j + 1:, :] -= np.outer(L[j + 1:, j], U[j, :])
        U[j + 1:, j] = 0.0

    def ab(A):
        return np.abs(A.real) + np.abs(A.imag)

    m = int(np.max(np.diff(plan.row_ptr)))
    return guard_bound_coeff(m) * np.max(ab(L) @ ab(U).sum(1)) / np.max(ab(M).sum(1))


# ------------------------------------------------------------- compact plan
LIN_HDR = ["n", "nnz", "nlev", "nblev", "ndyn", "ncon", "nsrc", "words", "o_lvl_ptr", "o_fdesc", "o_blvl_ptr",
           "o_bdesc", "o_col", "o_cpos", "o_rcon_ptr", "o_rcon", "o_dyn", "o_src", "lev0_free", "words_fwd",
           "ntlev", "ntblev", "o_tlvl_ptr", "o_tfdesc", "o_tblvl_ptr", "o_tbdesc", "o_tent", "nadj", "o_adj",
           "tlev0_free", "ne"]
LIN_HDR_WORDS = 64  # sizeof(LinHdr) / 2
LIN_DYN, LIN_SRC, LIN_ADJ, LIN_NONE = 4, 2, 8, 0xFFFF
LK_SRC, LK_CP, LK_CM, LK_IND = 0, 1, 2, 3
DYN_C, DYN_L, KG_V, KG_I = 1, 2, 3, 4
LIN = 3  # kg_plan_export kind of the compact plan


class LinPlan:
    """The compact 16-bit plan of the LDS-resident linear transient kernel
    (kg_plan_export kind 3), split into its arrays."""

    def __init__(self, elements, n, params, s_rep):
        el = np.ascontiguousarray(elements, _lib.ELEMENT_DTYPE)
        pr = np.ascontiguousarray(params, np.float64)
        L = _lib.lib()
        size = L.kg_plan_export(el.ctypes.data, len(el), n, LIN, NESTED, pr.ctypes.data, s_rep, None, 0, None, 0)
        self.status = -size if size < 0 else 0
        if size < 0:
            return
        buf = np.zeros(size, np.int32)
        L.kg_plan_export(el.ctypes.data, len(el), n, LIN, NESTED, pr.ctypes.data, s_rep, buf.ctypes.data, size,
                         None, 0)
        self.w = buf.view(np.uint16).astype(np.int64)
        self.h = {k: int(v) for k, v in zip(LIN_HDR, buf[:len(LIN_HDR)])}
        h = self.h
        assert h["words"] == len(self.w) and h["words"] % 8 == 0

        def a(name, k, align=2):
            o = h["o_" + name]
            assert LIN_HDR_WORDS <= o and o + k <= h["words"] and o % align == 0, name
            return self.w[o:o + k]
        n, nnz = h["n"], h["nnz"]
        self.lvl_ptr, self.blvl_ptr = a("lvl_ptr", h["nlev"] + 1), a("blvl_ptr", h["nblev"] + 1)
        self.fdesc = a("fdesc", 4 * n, 4).reshape(-1, 4)
        self.bdesc = a("bdesc", 4 * n, 4).reshape(-1, 4)
        self.col, self.cpos = a("col", nnz), a("cpos", n)
        self.rcon_ptr, self.rcon = a("rcon_ptr", n + 1), a("rcon", h["ncon"])
        self.dyn = a("dyn", LIN_DYN * h["ndyn"]).reshape(-1, LIN_DYN)
        self.src = a("src", LIN_SRC * h["nsrc"]).reshape(-1, LIN_SRC)
        self.dkind = self.dyn[:, 0] >> 14
        self.delem = self.dyn[:, 0] & 0x3FFF
        assert h["words_fwd"] <= h["words"] and h["words_fwd"] % 8 == 0
        for name in ("lvl_ptr", "fdesc", "blvl_ptr", "bdesc", "col", "cpos", "rcon_ptr", "rcon", "dyn", "src"):
            assert h["o_" + name] < h["words_fwd"], name  # the forward kernels copy only this prefix
        self.tlvl_ptr, self.tblvl_ptr = a("tlvl_ptr", h["ntlev"] + 1), a("tblvl_ptr", h["ntblev"] + 1)
        self.tfdesc = a("tfdesc", 4 * n, 4).reshape(-1, 4)
        self.tbdesc = a("tbdesc", 4 * n, 4).reshape(-1, 4)
        nt = int(max(self.tfdesc[:, 2].max(initial=0), self.tbdesc[:, 2].max(initial=0)))
        self.tent = a("tent", 2 * nt).reshape(-1, 2)
        self.adj = a("adj", LIN_ADJ * h["nadj"]).reshape(-1, LIN_ADJ)

    def check_indices(self, ne, full=None):
        """Every index the kernel dereferences, against the buffer it indexes;
        with the full plan, the descriptors against its rows."""
        h, n, nnz = self.h, self.h["n"], self.h["nnz"]
        for ptr, desc, nl in ((self.lvl_ptr, self.fdesc, h["nlev"]), (self.blvl_ptr, self.bdesc, h["nblev"])):
            assert ptr[0] == 0 and ptr[nl] == n and np.all(np.diff(ptr) >= 0)
            assert sorted(desc[:, 0].tolist()) == list(range(n))
            assert np.all(desc[:, 1] <= desc[:, 2]) and np.all(desc[:, 2] <= nnz) and np.all(desc[:, 3] < nnz)
        if full is not None:
            rp, dg = full.row_ptr, full.diag
            for i, p0, p1, d in self.fdesc:
                assert (p0, p1, d) == (rp[i], dg[i], dg[i])
            for i, p0, p1, d in self.bdesc:
                assert (p0, p1, d) == (dg[i] + 1, rp[i + 1], dg[i])
        assert np.all(self.col < n) and sorted(self.cpos.tolist()) == list(range(n))
        assert self.rcon_ptr[0] == 0 and self.rcon_ptr[n] == h["ncon"] and np.all(np.diff(self.rcon_ptr) >= 0)
        for code in self.rcon:
            k, idx = int(code) & 3, int(code) >> 2
            assert idx < (h["nsrc"] if k == LK_SRC else h["ndyn"])
            if k == LK_IND:
                assert self.dkind[idx] == DYN_L
            elif k != LK_SRC:
                assert self.dkind[idx] == DYN_C
        for j, r in enumerate(self.dyn):
            assert self.dkind[j] in (DYN_C, DYN_L) and self.delem[j] < ne
            assert all(v == LIN_NONE or v < n for v in r[1:4])
            assert self.dkind[j] != DYN_L or r[3] < n
        for r in self.src:
            assert r[0] in (KG_V, KG_I) and r[1] < ne
        if h["lev0_free"]:
            for i, p0, p1, _ in self.fdesc[self.lvl_ptr[0]:self.lvl_ptr[1]]:
                assert p0 == p1
        # adjoint extension
        for ptr, desc, nl in ((self.tlvl_ptr, self.tfdesc, h["ntlev"]), (self.tblvl_ptr, self.tbdesc, h["ntblev"])):
            assert ptr[0] == 0 and ptr[nl] == n and np.all(np.diff(ptr) >= 0)
            assert sorted(desc[:, 0].tolist()) == list(range(n))
            assert np.all(desc[:, 1] <= desc[:, 2]) and np.all(desc[:, 2] <= len(self.tent))
        assert np.all(self.tfdesc[:, 3] < nnz)
        assert np.all(self.tent[:, 0] < n) and np.all(self.tent[:, 1] < nnz)
        if h["tlev0_free"]:
            for ptr, desc in ((self.tlvl_ptr, self.tfdesc), (self.tblvl_ptr, self.tbdesc)):
                for _, q0, q1, _ in desc[ptr[0]:ptr[1]]:
                    assert q0 == q1
        for r in self.adj:
            kind, e = r[0] >> 14, r[0] & 0x3FFF
            assert kind in (0, 1, 2) and e < ne == h["ne"]
            assert all(x == LIN_NONE or x < n for x in r[1:7])
            if kind == 2:
                assert r[5] < n and r[6] < n and r[7] < h["ndyn"] and self.dkind[r[7]] == DYN_L
            if kind == 1:
                assert r[7] < h["ndyn"] and self.dkind[r[7]] == DYN_C


def wave_value(p, t):
    """The device's PULSE / SIN evaluation (voltage_source.rs:50-93)."""
    kind, w = int(p[2]), p[3:]
    if kind == 1:
        v1, v2, td, tr, tf, pw, per = w[:7]
        if t < td:
            return v1
        tt = math.fmod(t - td, per)
        if tt < tr:
            return v1 + (v2 - v1) * (tt / tr)
        if tt < tr + pw:
            return v2
        if tt < tr + pw + tf:
            return v2 + (v1 - v2) * ((tt - tr - pw) / tf)
        return v1
    if kind == 2:
        vo, va, f, td, th, ph = w[:6]
        if t < td:
            return vo
        return vo + va * math.exp(-th * (t - td)) * math.sin(2.0 * math.pi * f * (t - td) + ph * (math.pi / 180.0))
    return p[0]


def tran_lin(lp, vf, params, x0, h, nsteps, method):
    """The step loop of the LDS-resident linear transient kernel on the CPU:
    vf are the factored values (reciprocal diagonal), x0 the operating point
    in krets numbering. Returns [nsteps + 1][n] in krets numbering."""
    n = lp.h["n"]
    alpha = 2.0 if method == 2 else 1.0
    compat, trap = method == 0, method == 2
    vals = np.array([params[int(e), 0] for e in lp.delem])
    cf = alpha * (vals / h)
    sp = [params[int(r[1])] for r in lp.src]
    xq = np.zeros(n)
    xq[lp.cpos] = x0
    out = np.zeros((nsteps + 1, n))
    out[0] = x0

    def vd(x, r):
        return (x[r[1]] if r[1] != LIN_NONE else 0.0) - (x[r[2]] if r[2] != LIN_NONE else 0.0)
    ed = np.array([vd(xq, r) if k == DYN_L else 0.0 for r, k in zip(lp.dyn, lp.dkind)])
    lev0 = 1 if lp.h["lev0_free"] else 0
    for s in range(1, nsteps + 1):
        tsrc = (s - 1) * h if compat else s * h
        xc = np.zeros(n)
        for i in range(n):
            acc = 0.0
            for code in lp.rcon[lp.rcon_ptr[i]:lp.rcon_ptr[i + 1]]:
                k, idx = int(code) & 3, int(code) >> 2
                if k == LK_SRC:
                    acc += wave_value(sp[idx], tsrc) if lp.src[idx, 0] == KG_V else sp[idx][0]
                elif k == LK_IND:
                    rr = -cf[idx] * xq[lp.dyn[idx, 3]]
                    acc += rr - ed[idx] if alpha != 1.0 else rr
                else:
                    ieq = -cf[idx] * vd(xq, lp.dyn[idx])
                    if alpha != 1.0:
                        ieq -= ed[idx]
                    acc += -ieq if k == LK_CP else ieq
            xc[i] = acc
        for l in range(lev0, lp.h["nlev"]):
            for i, p0, p1, _ in lp.fdesc[lp.lvl_ptr[l]:lp.lvl_ptr[l + 1]]:
                xc[i] -= sum(vf[p] * xc[lp.col[p]] for p in range(p0, p1))
        for l in range(lp.h["nblev"]):
            for i, p0, p1, d in lp.bdesc[lp.blvl_ptr[l]:lp.blvl_ptr[l + 1]]:
                xc[i] = (xc[i] - sum(vf[p] * xc[lp.col[p]] for p in range(p0, p1))) * vf[d]
        if trap:
            for j, r in enumerate(lp.dyn):
                vn = vd(xc, r)
                ed[j] = cf[j] * (vn - vd(xq, r)) - ed[j] if lp.dkind[j] == DYN_C else vn
        out[s] = xc[lp.cpos]
        xq = xc
    return out


def _lin_setup(lp, vf, params, h, method):
    n = lp.h["n"]
    alpha = 2.0 if method == 2 else 1.0
    vals = np.array([params[int(e), 0] for e in lp.delem])
    return n, alpha, alpha * (vals / h)


def adj_lin(lp, vf, params, x0, h, nsteps, method, cvec, K):
    """The LDS-resident adjoint (forward with checkpoints, then the backward
    segments) on the CPU, in the kernels' order of operations; x0 and cvec
    in krets numbering. Returns (J, grad per element without the
    operating-point term, xbar0 in krets numbering)."""
    n, alpha, cf = _lin_setup(lp, vf, params, h, method)
    trap = method == 2
    NN = n

    def node(c):
        return NN if c == LIN_NONE else int(c)

    def step(xq, ed, s):
        x = np.zeros(n + 1)
        tsrc = s * h
        for i in range(n):
            acc = 0.0
            for code in lp.rcon[lp.rcon_ptr[i]:lp.rcon_ptr[i + 1]]:
                k, idx = int(code) & 3, int(code) >> 2
                if k == LK_SRC:
                    p = params[int(lp.src[idx, 1])]
                    acc += wave_value(p, tsrc) if lp.src[idx, 0] == KG_V else p[0]
                elif k == LK_IND:
                    rr = -cf[idx] * (xq[node(lp.dyn[idx, 3])] - 0.0)
                    acc += rr - ed[idx] if alpha != 1.0 else rr
                else:
                    r = lp.dyn[idx]
                    ieq = -cf[idx] * (xq[node(r[1])] - xq[node(r[2])])
                    if alpha != 1.0:
                        ieq -= ed[idx]
                    acc += -ieq if k == LK_CP else ieq
            x[i] = acc
        for l in range(lp.h["nlev"]):
            for i, p0, p1, _ in lp.fdesc[lp.lvl_ptr[l]:lp.lvl_ptr[l + 1]]:
                x[i] -= sum(vf[p] * x[lp.col[p]] for p in range(p0, p1))
        for l in range(lp.h["nblev"]):
            for i, p0, p1, d in lp.bdesc[lp.blvl_ptr[l]:lp.blvl_ptr[l + 1]]:
                x[i] = (x[i] - sum(vf[p] * x[lp.col[p]] for p in range(p0, p1))) * vf[d]
        ed = ed.copy()
        if trap:
            for j, r in enumerate(lp.dyn):
                vn = x[node(r[1])] - x[node(r[2])]
                ed[j] = cf[j] * (vn - (xq[node(r[1])] - xq[node(r[2])])) - ed[j] if lp.dkind[j] == DYN_C else vn
        return x, ed

    # forward with checkpoints
    xq = np.zeros(n + 1)
    xq[lp.cpos] = x0
    ed = np.array([xq[node(r[1])] - xq[node(r[2])] if k == DYN_L else 0.0 for r, k in zip(lp.dyn, lp.dkind)])
    ck = {}
    J = 0.0
    for s in range(nsteps + 1):
        if s > 0:
            xq, ed = step(xq, ed, s)
        w = 0.5 * h if s in (0, nsteps) else h
        J += w * float(np.dot(cvec, xq[lp.cpos]))
        if s % K == 0:
            ck[s // K] = (xq.copy(), ed.copy())
    # backward
    ne = lp.h["ne"]
    xbar = np.zeros(n + 1)
    xb0 = np.zeros(n + 1)
    xbar[lp.cpos] = 0.5 * h * cvec
    sbar = np.zeros(lp.h["ndyn"] + 1)
    gr = np.zeros(ne)
    nseg = (nsteps + K - 1) // K
    for sgi in range(nseg - 1, -1, -1):
        s0, s1 = sgi * K, min(sgi * K + K, nsteps)
        xs = [ck[sgi][0]]
        e_ = ck[sgi][1]
        for s in range(s0 + 1, s1 + 1):
            x, e_ = step(xs[-1], e_, s)
            xs.append(x)
        for s in range(s1 - 1, s0 - 1, -1):
            X1, Xs = xs[s + 1 - s0], xs[s - s0]
            if trap:
                for d, r in enumerate(lp.dyn):
                    sb1 = sbar[d]
                    if sb1 == 0.0:
                        continue
                    p, m = node(r[1]), node(r[2])
                    if lp.dkind[d] == DYN_C:
                        a_ = cf[d] * sb1
                        xbar[p] += a_
                        xb0[p] -= a_
                        xbar[m] -= a_
                        xb0[m] += a_
                        gr[lp.delem[d]] += sb1 * 2.0 / h * ((X1[p] - X1[m]) - (Xs[p] - Xs[m]))
                    else:
                        xbar[p] += sb1
                        xbar[m] -= sb1
            lam = np.zeros(n + 1)
            lam[:n] = xbar[:n]
            for l in range(lp.h["ntlev"]):
                for i, q0, q1, d in lp.tfdesc[lp.tlvl_ptr[l]:lp.tlvl_ptr[l + 1]]:
                    sm = lam[i] - sum(vf[lp.tent[q, 1]] * lam[lp.tent[q, 0]] for q in range(q0, q1))
                    lam[i] = sm * vf[d]
            for l in range(lp.h["ntblev"]):
                for i, q0, q1, _ in lp.tbdesc[lp.tblvl_ptr[l]:lp.tblvl_ptr[l + 1]]:
                    lam[i] -= sum(vf[lp.tent[q, 1]] * lam[lp.tent[q, 0]] for q in range(q0, q1))
            for r in lp.adj:
                kind, e = int(r[0]) >> 14, int(r[0]) & 0x3FFF
                v = params[e, 0]
                p, m = node(r[1]), node(r[2])
                if kind == 2:
                    c, rc, d = int(r[5]), int(r[6]), int(r[7])
                    lb = lam[rc]
                    xb0[c] += -alpha * (v / h) * lb
                    sbar[d] = -lb if trap else 0.0
                    gr[e] += lb * (-alpha / h) * Xs[c] + alpha / h * lb * X1[c]
                    continue
                el = lam[node(r[3])] - lam[node(r[4])]
                d1 = X1[p] - X1[m]
                if kind == 0:
                    gr[e] += el * d1 / (v * v)
                else:
                    d = int(r[7])
                    d0 = Xs[p] - Xs[m]
                    aa = alpha * (v / h) * el
                    xb0[p] += aa
                    xb0[m] -= aa
                    sbar[d] = (-sbar[d] + el) if trap else 0.0
                    gr[e] += el * alpha / h * d0 - alpha / h * el * d1
            w = 0.5 * h if s == 0 else h
            for g in range(n):
                j = lp.cpos[g]
                xbar[j] = xb0[j] + w * cvec[g]
                xb0[j] = 0.0
            xb0[n] = xbar[n] = 0.0
    if trap:
        for d, r in enumerate(lp.dyn):
            if lp.dkind[d] == DYN_L:
                xbar[node(r[1])] += sbar[d]
                xbar[node(r[2])] -= sbar[d]
    return J, gr, xbar[lp.cpos]
