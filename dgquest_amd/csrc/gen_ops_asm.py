#!/usr/bin/env python3
"""Generates ops_asm.inc: the op interpreter of the sweep kernel in gfx950
assembly (one inline-asm body per amplitude type and register bits R) and
the op-code table the engine's host code uses to build device programs.

Why assembly: the register tile of a thread (2^R complex amplitudes) must be
updated in place by a loop over a runtime op list.  Compiled from C++, the
loop's switch over (op kind, register bit) makes the register allocator keep
two copies of the tile and copy between them on every op (and spill at R=4),
which costs more VALU issue than the arithmetic itself.  Inside one asm
statement every tile value is a tied "+v" operand, so it lives in one
register for the whole op list.

Why few scalar instructions: a CU issues one scalar instruction per cycle for
all its waves, as many as its four SIMDs issue 64-lane vector instructions,
so per-op bookkeeping (decoding, branches, pointer updates) competes with the
arithmetic.  Ops are dispatched through a jump table, have fixed sizes per op
code, end the list with an END op, carry workgroup conditions in a separate
WG op only when they have one, and come in variants whose register condition
is resolved at generation time (no per-register branches).

Device program (built by the engine's host code from dq_op arrays): each op is
a 16-byte header followed by its coefficients in the amplitude type; the
size of each op code is fixed (op_size below).
    u32 code     index into the jump table
    u32 mask     SHEARS / TSH_*: register bits sheared; *_GEN / MAT / SWAP:
                 register indices (pair ops: lower indices) that satisfy the
                 condition
    u32 c_thr    lane condition mask (0: every lane); TSH_GEN: the table's
                 register indices ...
    u32 v_thr    ... and value; WG: bytes of the guarded op
    coefficients from byte 16:
      WG        c_wg (u64), v_wg (u64)
      SHEARS    per register bit b: u_b, l_b at 16 + 2 * b * sizeof(T)
      TABLE_*   2^R factors (HALF: the 2^(R-1) with the bit set, in order)
      TSH_*     a TABLE_* table, then the SHEARS slots: the diagonal, then the
                shears, in one op
      PHASE_*   re, im          MAT  u00 u01 u10 u11 (re, im)     RMAT m0..m3
Scalar registers s36..s71 are reserved by the asm (clobbered; the compiler
keeps the SGPRs above s71 for itself, and at 8 waves per SIMD it may not use
any above s71):
    s[36:51] second load buffer / scratch (s[36:37]: saved exec)
    s[52:67] header + first coefficients     s68 byte offset of the current op
    s69      saved mask                       s[70:71] jump-table address
Each handler ends by loading the next op's header (program base + s68) and
jumping through the table, with no branch back to a common loop head.
Only scalar loads touch memory; every store is a vector register write.

Usage: python3 gen_ops_asm.py > ops_asm.inc   (the Makefile does this)
"""
import itertools
import sys

BUF2, HDR, OFF, SCR, TAB = 36, 52, 68, 69, 70
EXEC = BUF2             # saved exec of a lane-conditioned op (s[36:37])
COEF = HDR + 4          # coefficients start right after the 16-byte header
CLOBBER = range(36, 72)
RMAX = 5


def sp(b):
    return f"s[{b}:{b + 1}]"


def op_codes(R):
    """Op codes, the same list for every amplitude type: (name, params)."""
    codes = [("END", ()), ("WG", ()), ("SHEARS", ()), ("TABLE_FULL", ()), ("TABLE_GEN", ())]
    codes += [("TABLE_HALF", (r,)) for r in range(R)]
    for k in ("NEG", "PHASE"):
        codes += [(f"{k}_ALL", ()), (f"{k}_GEN", ())]
        codes += [(f"{k}_B", (r, v)) for r in range(R) for v in (0, 1)]
        codes += [(f"{k}_BB", (r1, r2, v1, v2)) for r1, r2 in itertools.combinations(range(R), 2)
                  for v1 in (0, 1) for v2 in (0, 1)]
    codes += [("MAT", (b,)) for b in range(R)]
    codes += [("SWAP", (b,)) for b in range(R)]
    codes += [("RMAT", (b,)) for b in range(R)]
    codes += [("TSH_FULL", ()), ("TSH_GEN", ())]
    codes += [("TSH_HALF", (r,)) for r in range(R)]
    return codes


def op_size(name, R, ts):
    """Bytes of an op (a multiple of 16)."""
    base = name.split("_")[0]
    if name in ("END", "SWAP") or base == "NEG":
        b = 0
    elif name == "WG":
        b = 16
    elif name == "SHEARS":
        b = 2 * R * ts
    elif name in ("TABLE_FULL", "TABLE_GEN"):
        b = (2 << R) * ts
    elif name == "TABLE_HALF":
        b = (1 << R) * ts
    elif name in ("TSH_FULL", "TSH_GEN"):
        b = (2 << R) * ts + 2 * R * ts
    elif name == "TSH_HALF":
        b = (1 << R) * ts + 2 * R * ts
    elif base == "PHASE":
        b = 2 * ts
    elif name == "MAT":
        b = 8 * ts
    elif name == "RMAT":
        b = 4 * ts
    else:
        raise ValueError(name)
    return (16 + b + 15) // 16 * 16


class Gen:
    def __init__(self, T, R):
        self.T, self.R, self.NR = T, R, 1 << R
        self.d = T == "double"
        self.ts = 8 if self.d else 4
        self.L = []
        self.codes = op_codes(R)

    def e(self, s):
        self.L.append(s)

    def re(self, j):
        return f"%[re{j}]"

    def im(self, j):
        return f"%[im{j}]"

    def sval(self, base, i):
        """Value i (in units of T) of a coefficient buffer starting at s<base>."""
        return sp(base + 2 * i) if self.d else f"s{base + i}"

    @property
    def f(self):
        return "f64" if self.d else "f32"

    @property
    def mov(self):
        return "v_mov_b64" if self.d else "v_mov_b32"

    def lbl(self, name):
        return f".L{name}_%="

    def dispatch(self):
        """load the header of the op at s68 and jump: table entry = s72 + 4 * code"""
        self.e(f"s_load_dwordx16 s[{HDR}:{HDR + 15}], %[prog], s{OFF}")
        self.e("s_waitcnt lgkmcnt(0)")
        self.e(f"s_lshl2_add_u32 s36, s{HDR}, s{TAB}")
        self.e(f"s_addc_u32 s37, s{TAB + 1}, 0")
        self.e("s_setpc_b64 s[36:37]")

    def tail(self, name):
        size = op_size(name, self.R, self.ts)
        self.e(f"s_add_u32 s{OFF}, s{OFF}, {size}")
        self.dispatch()

    def load(self, dst, off):
        """64 bytes of the current op from byte `off` into s[dst:dst+15]"""
        self.e(f"s_load_dwordx16 s[{dst}:{dst + 15}], %[prog], s{OFF} offset:{hex(off)}")

    def load_bytes(self, off, nbytes):
        """nbytes (a multiple of 4, at most 128) of the current op from byte
        `off` into BUF2, then HDR; returns the first SGPR of each dword"""
        regs = list(range(BUF2, BUF2 + 16)) + list(range(HDR, HDR + 16))
        nd = nbytes // 4
        i = 0
        while i < nd:
            for w in (16, 8, 4, 2, 1):
                # a load stays inside one buffer and starts at a multiple of min(w, 4)
                if w <= nd - i and (i % 16) + w <= 16 and regs[i] % min(w, 4) == 0:
                    break
            op = "s_load_dword" if w == 1 else f"s_load_dwordx{w}"
            dst = f"s{regs[i]}" if w == 1 else f"s[{regs[i]}:{regs[i] + w - 1}]"
            self.e(f"{op} {dst}, %[prog], s{OFF} offset:{hex(off + 4 * i)}")
            i += w
        return regs[:nd]

    # --- arithmetic building blocks ------------------------------------------
    def pairs(self, B):
        return [(j0, j0 | (1 << B)) for j0 in range(self.NR) if not (j0 >> B) & 1]

    def cmul(self, j, pr, pi):
        f = self.f
        self.e(f"v_mul_{f} %[t0], {pi}, {self.im(j)}")
        self.e(f"v_mul_{f} %[t1], {pi}, {self.re(j)}")
        self.e(f"v_fma_{f} {self.re(j)}, {pr}, {self.re(j)}, -%[t0]")
        self.e(f"v_fma_{f} {self.im(j)}, {pr}, {self.im(j)}, %[t1]")

    def cmul_list(self, js, pr_of, pi_of):
        """two amplitudes at a time (temporaries t0..t3) so products overlap."""
        f = self.f
        for k in range(0, len(js), 2):
            grp = js[k:k + 2]
            for n, j in enumerate(grp):
                self.e(f"v_mul_{f} %[t{2 * n}], {pi_of(j)}, {self.im(j)}")
                self.e(f"v_mul_{f} %[t{2 * n + 1}], {pi_of(j)}, {self.re(j)}")
            for n, j in enumerate(grp):
                self.e(f"v_fma_{f} {self.re(j)}, {pr_of(j)}, {self.re(j)}, -%[t{2 * n}]")
                self.e(f"v_fma_{f} {self.im(j)}, {pr_of(j)}, {self.im(j)}, %[t{2 * n + 1}]")

    def lanes_on(self, tag):
        """exec &= lanes with (tid & c_thr) == v_thr; exec saved in s[36:37]."""
        self.e(f"s_mov_b64 {sp(EXEC)}, exec")
        self.e(f"s_cmp_eq_u32 s{HDR + 2}, 0")
        self.e(f"s_cbranch_scc1 {self.lbl(tag)}")
        self.e(f"v_and_b32 %[vt], s{HDR + 2}, %[tid]")
        self.e(f"v_cmp_eq_u32 vcc, s{HDR + 3}, %[vt]")
        self.e("s_and_b64 exec, exec, vcc")
        self.e(f"{self.lbl(tag)}:")

    def lanes_off(self):
        self.e(f"s_mov_b64 exec, {sp(EXEC)}")

    def jskip(self, j, tag):
        self.e(f"s_bitcmp1_b32 s{HDR + 1}, {j}")
        self.e(f"s_cbranch_scc0 {self.lbl(tag)}")

    # --- table factors (starting at op + 16) in 64-byte chunks alternating
    # between s[52:67] and s[36:51]; scalar loads return out of order, so
    # every wait is lgkmcnt(0) -----------------------------------------------------
    def table_apply(self, entries):
        """entries: (register index j, position of its factor in the stored
        table).  Loads only the chunks that hold a needed factor."""
        per = 4 if self.d else 8
        chunks = sorted({p // per for _, p in entries})
        bufs = [HDR, BUF2]
        where = {}
        for i, c in enumerate(chunks[:2]):
            self.load(bufs[i], 16 + 64 * c)
            where[c] = bufs[i]
        self.e("s_waitcnt lgkmcnt(0)")
        for i, c in enumerate(chunks):
            buf = where[c]
            pos = {j: p for j, p in entries if p // per == c}
            js = sorted(pos)
            self.cmul_list(js, lambda j, buf=buf, pos=pos: self.sval(buf, 2 * (pos[j] % per)),
                           lambda j, buf=buf, pos=pos: self.sval(buf, 2 * (pos[j] % per) + 1))
            if i + 2 < len(chunks):  # refill this buffer with the chunk after next
                c2 = chunks[i + 2]
                self.load(buf, 16 + 64 * c2)
                where[c2] = buf
            if i % 2 == 1 and i + 1 < len(chunks):
                self.e("s_waitcnt lgkmcnt(0)")

    # --- handlers ---------------------------------------------------------------
    def h_wg(self):
        # skip the guarded op (its size in the header) unless (base & c_wg) == v_wg
        self.e(f"s_and_b64 s[36:37], %[base], {sp(COEF)}")
        self.e(f"s_cmp_eq_u64 s[36:37], {sp(COEF + 2)}")
        self.e(f"s_cselect_b32 s38, 0, s{HDR + 3}")
        self.e(f"s_add_u32 s{OFF}, s{OFF}, s38")
        self.tail("WG")

    def shear_bits(self, mask_reg, coef, tag):
        """the shears of every register bit set in s<mask_reg>; coef(B) gives
        the (u, l) operands of bit B"""
        f = self.f
        for B in range(self.R):
            u, l = coef(B)
            t = f"{tag}{B}"
            self.e(f"s_bitcmp1_b32 s{mask_reg}, {B}")
            self.e(f"s_cbranch_scc0 {self.lbl(t)}")
            P = self.pairs(B)
            for j0, j1 in P:
                self.e(f"v_fmac_{f} {self.re(j0)}, {u}, {self.re(j1)}")
                self.e(f"v_fmac_{f} {self.im(j0)}, {u}, {self.im(j1)}")
            for j0, j1 in P:
                self.e(f"v_fmac_{f} {self.re(j1)}, {l}, {self.re(j0)}")
                self.e(f"v_fmac_{f} {self.im(j1)}, {l}, {self.im(j0)}")
            self.e(f"{self.lbl(t)}:")

    def h_shears(self):
        if self.d and self.R >= 4:  # slots 3 (and 4) beyond the first load
            self.e(f"s_load_dwordx8 s[{BUF2}:{BUF2 + 7}], %[prog], s{OFF} offset:0x40")
            self.e("s_waitcnt lgkmcnt(0)")

        def coef(B):
            if self.d:
                if B < 3:
                    return sp(COEF + 4 * B), sp(COEF + 4 * B + 2)
                return sp(BUF2 + 4 * (B - 3)), sp(BUF2 + 4 * (B - 3) + 2)
            return f"s{COEF + 2 * B}", f"s{COEF + 2 * B + 1}"

        self.shear_bits(HDR + 1, coef, "sh")
        self.tail("SHEARS")

    def h_tsh(self, name, params, tag):
        """a diagonal table, then the shears of the bits in the header mask"""
        ts = self.ts
        self.e(f"s_mov_b32 s{SCR}, s{HDR + 1}")        # shear mask
        if name == "TSH_GEN":
            self.e(f"s_mov_b32 s{BUF2}, s{HDR + 2}")   # table register indices (table_gen loads into HDR only)
        if name == "TSH_FULL":
            self.table_apply([(j, j) for j in range(self.NR)])
            tbytes = (2 << self.R) * ts
        elif name == "TSH_HALF":
            (r,) = params
            js = [j for j in range(self.NR) if (j >> r) & 1]
            self.table_apply([(j, i) for i, j in enumerate(js)])
            tbytes = (1 << self.R) * ts
        else:
            self.table_gen(BUF2, tag)
            tbytes = (2 << self.R) * ts
        regs = self.load_bytes(16 + tbytes, 2 * self.R * ts)
        self.e("s_waitcnt lgkmcnt(0)")

        def coef(B):
            if self.d:
                return sp(regs[4 * B]), sp(regs[4 * B + 2])
            return f"s{regs[2 * B]}", f"s{regs[2 * B + 1]}"

        self.shear_bits(SCR, coef, tag + "s")
        self.tail(name)

    def h_table(self, name, params):
        if name == "TABLE_FULL":
            self.table_apply([(j, j) for j in range(self.NR)])
        elif name == "TABLE_HALF":
            (r,) = params
            js = [j for j in range(self.NR) if (j >> r) & 1]
            self.table_apply([(j, i) for i, j in enumerate(js)])
        else:  # TABLE_GEN: per-register skip on the mask (kept in s69 while loading)
            self.e(f"s_mov_b32 s{SCR}, s{HDR + 1}")
            self.table_gen(SCR, "tg")
        self.tail(name)

    def table_gen(self, mask_reg, tag):
        """the table factors of the register indices set in s<mask_reg>"""
        per = 4 if self.d else 8
        for c in range((self.NR + per - 1) // per):
            self.load(HDR, 16 + 64 * c)
            self.e("s_waitcnt lgkmcnt(0)")
            for j in range(c * per, min(self.NR, (c + 1) * per)):
                t = f"{tag}g{j}"
                self.e(f"s_bitcmp1_b32 s{mask_reg}, {j}")
                self.e(f"s_cbranch_scc0 {self.lbl(t)}")
                self.cmul(j, self.sval(HDR, 2 * (j % per)), self.sval(HDR, 2 * (j % per) + 1))
                self.e(f"{self.lbl(t)}:")

    def sel(self, name, params):
        """register indices an ALL / B / BB variant acts on (None: GEN, the mask)."""
        kind = name.split("_")[1]
        if kind == "ALL":
            return list(range(self.NR))
        if kind == "B":
            r, v = params
            return [j for j in range(self.NR) if ((j >> r) & 1) == v]
        if kind == "BB":
            r1, r2, v1, v2 = params
            return [j for j in range(self.NR) if ((j >> r1) & 1) == v1 and ((j >> r2) & 1) == v2]
        return None

    def h_neg(self, name, params, tag):
        f = self.f
        self.lanes_on(tag + "l")
        js = self.sel(name, params)
        for j in (js if js is not None else range(self.NR)):
            if js is None:
                self.jskip(j, f"{tag}j{j}")
            self.e(f"v_mul_{f} {self.re(j)}, -1.0, {self.re(j)}")
            self.e(f"v_mul_{f} {self.im(j)}, -1.0, {self.im(j)}")
            if js is None:
                self.e(f"{self.lbl(f'{tag}j{j}')}:")
        self.lanes_off()
        self.tail(name)

    def h_phase(self, name, params, tag):
        self.lanes_on(tag + "l")
        pr, pi = self.sval(COEF, 0), self.sval(COEF, 1)
        js = self.sel(name, params)
        if js is not None:
            self.cmul_list(js, lambda j: pr, lambda j: pi)
        else:
            for j in range(self.NR):
                self.jskip(j, f"{tag}j{j}")
                self.cmul(j, pr, pi)
                self.e(f"{self.lbl(f'{tag}j{j}')}:")
        self.lanes_off()
        self.tail(name)

    def h_mat(self, params):
        (B,) = params
        f = self.f
        if self.d:  # u11 (re, im) lies beyond the first load (s[36:37] will hold the saved exec)
            self.e(f"s_load_dwordx4 s[{BUF2 + 4}:{BUF2 + 7}], %[prog], s{OFF} offset:0x40")
            u = [self.sval(COEF, i) for i in range(6)] + [sp(BUF2 + 4), sp(BUF2 + 6)]
        else:
            u = [self.sval(COEF, i) for i in range(8)]
        self.lanes_on(f"mt{B}l")
        if self.d:
            self.e("s_waitcnt lgkmcnt(0)")
        u0r, u0i, u1r, u1i, u2r, u2i, u3r, u3i = u
        for j0, j1 in self.pairs(B):
            tag = f"mt{B}_{j0}"
            self.jskip(j0, tag)
            xr, xi, yr, yi = self.re(j0), self.im(j0), self.re(j1), self.im(j1)
            self.e(f"v_mul_{f} %[t0], {u0r}, {xr}")
            self.e(f"v_mul_{f} %[t1], {u0r}, {xi}")
            self.e(f"v_mul_{f} %[t2], {u2r}, {xr}")
            self.e(f"v_mul_{f} %[t3], {u2r}, {xi}")
            self.e(f"v_fma_{f} %[t0], -{u0i}, {xi}, %[t0]")
            self.e(f"v_fma_{f} %[t1], {u0i}, {xr}, %[t1]")
            self.e(f"v_fma_{f} %[t2], -{u2i}, {xi}, %[t2]")
            self.e(f"v_fma_{f} %[t3], {u2i}, {xr}, %[t3]")
            self.e(f"v_fma_{f} %[t0], {u1r}, {yr}, %[t0]")
            self.e(f"v_fma_{f} %[t1], {u1r}, {yi}, %[t1]")
            self.e(f"v_fma_{f} %[t2], {u3r}, {yr}, %[t2]")
            self.e(f"v_fma_{f} %[t3], {u3i}, {yr}, %[t3]")
            self.e(f"v_fma_{f} %[t0], -{u1i}, {yi}, %[t0]")
            self.e(f"v_fma_{f} %[t1], {u1i}, {yr}, %[t1]")
            self.e(f"v_fma_{f} %[t2], -{u3i}, {yi}, %[t2]")
            self.e(f"v_fma_{f} {yi}, {u3r}, {yi}, %[t3]")
            self.e(f"{self.mov} {xr}, %[t0]")
            self.e(f"{self.mov} {xi}, %[t1]")
            self.e(f"{self.mov} {yr}, %[t2]")
            self.e(f"{self.lbl(tag)}:")
        self.lanes_off()
        self.tail("MAT")

    def h_swap(self, params):
        (B,) = params
        self.lanes_on(f"sw{B}l")
        for j0, j1 in self.pairs(B):
            tag = f"sw{B}_{j0}"
            self.jskip(j0, tag)
            for a, b in ((self.re(j0), self.re(j1)), (self.im(j0), self.im(j1))):
                self.e(f"{self.mov} %[t0], {a}")
                self.e(f"{self.mov} {a}, {b}")
                self.e(f"{self.mov} {b}, %[t0]")
            self.e(f"{self.lbl(tag)}:")
        self.lanes_off()
        self.tail("SWAP")

    def h_rmat(self, params):
        (B,) = params
        f = self.f
        m0, m1, m2, m3 = (self.sval(COEF, i) for i in range(4))
        for j0, j1 in self.pairs(B):
            xr, xi, yr, yi = self.re(j0), self.im(j0), self.re(j1), self.im(j1)
            self.e(f"v_mul_{f} %[t0], {m0}, {xr}")
            self.e(f"v_mul_{f} %[t1], {m0}, {xi}")
            self.e(f"v_mul_{f} %[t2], {m2}, {xr}")
            self.e(f"v_mul_{f} %[t3], {m2}, {xi}")
            self.e(f"v_fma_{f} %[t0], {m1}, {yr}, %[t0]")
            self.e(f"v_fma_{f} %[t1], {m1}, {yi}, %[t1]")
            self.e(f"v_fma_{f} {yr}, {m3}, {yr}, %[t2]")
            self.e(f"v_fma_{f} {yi}, {m3}, {yi}, %[t3]")
            self.e(f"{self.mov} {xr}, %[t0]")
            self.e(f"{self.mov} {xi}, %[t1]")
        self.tail("RMAT")

    # --- the whole interpreter -----------------------------------------------
    def build(self):
        e = self.e
        # jump table: entry = .Ltab + 4 * code (every entry is one s_branch)
        e(f"s_mov_b32 s{OFF}, 0")
        e(f"s_getpc_b64 s[{TAB}:{TAB + 1}]")
        e(f"{self.lbl('pc')}:")
        e(f"s_add_u32 s{TAB}, s{TAB}, ({self.lbl('tab')} - {self.lbl('pc')})")
        e(f"s_addc_u32 s{TAB + 1}, s{TAB + 1}, 0")
        self.dispatch()
        e(f"{self.lbl('tab')}:")
        for i in range(len(self.codes)):
            e(f"s_branch {self.lbl(f'h{i}')}")
        for i, (name, params) in enumerate(self.codes):
            e(f"{self.lbl(f'h{i}')}:")
            tag = f"c{i}"
            if name == "END":
                e(f"s_branch {self.lbl('end')}")
            elif name == "WG":
                self.h_wg()
            elif name == "SHEARS":
                self.h_shears()
            elif name.startswith("TABLE"):
                self.h_table(name, params)
            elif name.startswith("NEG"):
                self.h_neg(name, params, tag)
            elif name.startswith("PHASE"):
                self.h_phase(name, params, tag)
            elif name == "MAT":
                self.h_mat(params)
            elif name == "SWAP":
                self.h_swap(params)
            elif name == "RMAT":
                self.h_rmat(params)
            elif name.startswith("TSH"):
                self.h_tsh(name, params, tag)
        e(f"{self.lbl('end')}:")
        return self.L

    def emit(self):
        body = self.build()
        out = []
        out.append(f"template <> struct AsmOps<{self.T}, {self.R}> {{")
        out.append(f"    __device__ __forceinline__ static void run(Regs<{self.T}, {self.R}> &a, const void *prog,")
        out.append("                                               uint64_t base, uint32_t tid) {")
        out.append(f"        {self.T} t0, t1, t2, t3;")
        out.append("        uint32_t vt;")
        out.append("        asm volatile(")
        for ln in body:
            pre = "" if ln.endswith(":") else "\\t"
            out.append(f'            "{pre}{ln}\\n"')
        outs = [f'[re{j}] "+v"(a.re[{j}])' for j in range(self.NR)] + \
               [f'[im{j}] "+v"(a.im[{j}])' for j in range(self.NR)] + \
               ['[t0] "=&v"(t0)', '[t1] "=&v"(t1)', '[t2] "=&v"(t2)', '[t3] "=&v"(t3)', '[vt] "=&v"(vt)']
        ins = ['[prog] "s"(prog)', '[base] "s"(base)', '[tid] "v"(tid)']
        out.append("            : " + ",\n              ".join(outs))
        out.append("            : " + ", ".join(ins))
        clob = ", ".join(f'"s{i}"' for i in CLOBBER)
        out.append(f'            : {clob}, "vcc", "scc", "memory");')
        out.append("    }")
        out.append("};")
        return "\n".join(out)


def code_table():
    """C++ view of the op codes for the host-side program builder."""
    out = ["// op codes of the device program per register bits R (index = code)",
           "struct DevOpCode {",
           "    const char *name;",
           "    int p[4];",
           "    int size64, size32;  // bytes for double / float amplitudes",
           "};"]
    for R in range(1, RMAX + 1):
        out.append(f"static const DevOpCode kDevOps{R}[] = {{")
        for name, params in op_codes(R):
            p = list(params) + [-1] * (4 - len(params))
            out.append(f'    {{"{name}", {{{p[0]}, {p[1]}, {p[2]}, {p[3]}}}, {op_size(name, R, 8)}, '
                       f'{op_size(name, R, 4)}}},')
        out.append("};")
    out.append("static inline const DevOpCode *dev_ops(int R, int &count) {")
    out.append("    switch (R) {")
    for R in range(1, RMAX + 1):
        out.append(f"        case {R}: count = (int)(sizeof(kDevOps{R}) / sizeof(DevOpCode)); return kDevOps{R};")
    out.append("        default: count = 0; return nullptr;")
    out.append("    }")
    out.append("}")
    return "\n".join(out)


def main():
    print("// Generated by gen_ops_asm.py -- the sweep kernel's op interpreter in gfx950")
    print("// assembly, one body per (amplitude type, register bits), and the op-code")
    print("// table of its device programs.  Do not edit.")
    print(code_table())
    print()
    print("template <typename T, int R> struct AsmOps;")
    for T in ("double", "float"):
        for R in range(1, RMAX + 1):
            print(Gen(T, R).emit())
            print()


if __name__ == "__main__":
    sys.exit(main())
