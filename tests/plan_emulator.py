"""CPU emulation of the device algorithm on a symbolic plan (kg_plan_export).

Test infrastructure: it runs the same gather / factor-program / level-order
solve steps the gfx950 kernels run, in numpy, so a plan can be checked on a
machine without a GPU.
"""
import ctypes

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
