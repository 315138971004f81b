#!/usr/bin/env python3
"""Generates ops_asm.inc: the op interpreter of the sweep kernel in gfx950
assembly, one inline-asm body per (amplitude type, register bits R).

Why assembly: the register tile of a thread (2^R complex amplitudes) must be
updated in place by a loop over a runtime op list.  Compiled from C++, the
loop's switch over (op kind, register bit) makes the register allocator keep
two copies of the tile and copy between them on every op (and spill at R=4),
which costs more VALU issue than the arithmetic itself.  Inside one asm
statement every tile value is a tied "+v" operand, so it lives in one
register for the whole op list.

Device program format (built by the engine's host code from dq_op arrays):
each op is a 32-byte header followed by coefficients of the amplitude type,
padded to a multiple of 32 bytes:
    u32 disp    kind * 8 + register bit
    u32 jmask   register indices j that satisfy the op's register condition
                (pair ops: the lower index of each pair; DIAGR: entries != 1)
    u32 thr     lane condition: c_thr | v_thr << 16 (c_thr = 0: every lane)
    u32 size    bytes of this op, header included
    u64 c_wg    workgroup condition: (tile base & c_wg) == v_wg
    u64 v_wg
    T   coef[]  SHEAR u, l | PHASE re, im | RMAT m0..m3 | MAT u00..u11 (re, im)
                | DIAGR 2^R complex factors | NEG, SWAP: none
Kinds (device numbering): 1 SHEAR, 2 DIAGR, 3 NEG, 4 PHASE, 5 MAT, 6 SWAP, 7 RMAT.

Scalar registers s36..s68 are reserved by the asm (clobbered; the compiler
keeps the SGPRs above s71 for itself at high occupancy):
    s[36:37] scratch    s[38:39] saved exec  s[40:41] next op  s[42:43] current op
    s[44:51] second coefficient buffer (FP64 MAT u10, u11)
    s[52:67] header + first 32 bytes of coefficients (also DIAGR table chunks)
    s68      ops left
Only scalar loads touch memory; every store is a vector register write.

Usage: python3 gen_ops_asm.py > ops_asm.inc   (the Makefile does this)
"""
import sys

KIND = {"SHEAR": 1, "DIAGR": 2, "NEG": 3, "PHASE": 4, "MAT": 5, "SWAP": 6, "RMAT": 7}

# scalar register map (see the module docstring)
SCR, EXEC, NEXT, CUR, BUF2, HDR, CNT = 36, 38, 40, 42, 44, 52, 68
COEF = HDR + 8          # first coefficient right after the 32-byte header
CLOBBER = range(36, 69)


def sp(b):
    return f"s[{b}:{b + 1}]"


class Gen:
    def __init__(self, T, R):
        self.T, self.R, self.NR = T, R, 1 << R
        self.d = T == "double"
        self.L = []

    def e(self, s):
        self.L.append(s)

    # --- operand and register names ----------------------------------------
    def re(self, j):
        return f"%[re{j}]"

    def im(self, j):
        return f"%[im{j}]"

    def sbuf(self, base, i):
        """Value i (in units of T) of a coefficient buffer starting at s<base>."""
        return sp(base + 2 * i) if self.d else f"s{base + i}"

    def sreg(self, i):
        return self.sbuf(COEF, i)

    @property
    def sfx(self):
        return "f64" if self.d else "f32"

    @property
    def mov(self):
        return "v_mov_b64" if self.d else "v_mov_b32"

    def lbl(self, name):
        return f".L{name}_%="

    # --- op bodies ------------------------------------------------------------
    def pairs(self, B):
        return [(j0, j0 | (1 << B)) for j0 in range(self.NR) if not (j0 >> B) & 1]

    def shear(self, B):
        f = self.sfx
        u, l = self.sreg(0), self.sreg(1)
        P = self.pairs(B)
        for j0, j1 in P:
            self.e(f"v_fmac_{f} {self.re(j0)}, {u}, {self.re(j1)}")
            self.e(f"v_fmac_{f} {self.im(j0)}, {u}, {self.im(j1)}")
        for j0, j1 in P:
            self.e(f"v_fmac_{f} {self.re(j1)}, {l}, {self.re(j0)}")
            self.e(f"v_fmac_{f} {self.im(j1)}, {l}, {self.im(j0)}")

    def cmul(self, js, pr_of, pi_of):
        """amplitudes js *= (pr + i pi), two at a time (temporaries t0..t3)."""
        f = self.sfx
        for k in range(0, len(js), 2):
            grp = js[k:k + 2]
            for n, j in enumerate(grp):
                pi = pi_of(j)
                self.e(f"v_mul_{f} %[t{2 * n}], {pi}, {self.im(j)}")
                self.e(f"v_mul_{f} %[t{2 * n + 1}], {pi}, {self.re(j)}")
            for n, j in enumerate(grp):
                pr = pr_of(j)
                self.e(f"v_fma_{f} {self.re(j)}, {pr}, {self.re(j)}, -%[t{2 * n}]")
                self.e(f"v_fma_{f} {self.im(j)}, {pr}, {self.im(j)}, %[t{2 * n + 1}]")

    def diagr(self):
        # table at current op + 32, read in 64-byte chunks into the header
        # buffer (the header is no longer needed once the op is dispatched)
        # (jmask, moved to s[SCR] before the header buffer is reused, skips the
        # registers whose factor is exactly 1)
        per = 4 if self.d else 8          # complex entries per 64-byte chunk
        nch = (self.NR + per - 1) // per
        f = self.sfx
        self.e(f"s_mov_b32 s{SCR}, s{HDR + 1}")
        for c in range(nch):
            self.e(f"s_load_dwordx16 s[{HDR}:{HDR + 15}], {sp(CUR)}, {hex(0x20 + 64 * c)}")
            self.e("s_waitcnt lgkmcnt(0)")
            for j in range(c * per, min(self.NR, (c + 1) * per)):
                pr, pi = self.sbuf(HDR, 2 * (j % per)), self.sbuf(HDR, 2 * (j % per) + 1)
                tag = f"dg{j}"
                self.e(f"s_bitcmp1_b32 s{SCR}, {j}")
                self.e(f"s_cbranch_scc0 {self.lbl(tag)}")
                self.e(f"v_mul_{f} %[t0], {pi}, {self.im(j)}")
                self.e(f"v_mul_{f} %[t1], {pi}, {self.re(j)}")
                self.e(f"v_fma_{f} {self.re(j)}, {pr}, {self.re(j)}, -%[t0]")
                self.e(f"v_fma_{f} {self.im(j)}, {pr}, {self.im(j)}, %[t1]")
                self.e(f"{self.lbl(tag)}:")

    def jskip(self, j, tag):
        self.e(f"s_bitcmp1_b32 s{HDR + 1}, {j}")
        self.e(f"s_cbranch_scc0 {self.lbl(tag)}")

    def lanes_on(self, tag):
        """exec &= lanes satisfying (tid & c_thr) == v_thr (exec is saved before
        dispatch and restored after every op)."""
        self.e(f"s_and_b32 s{SCR}, s{HDR + 2}, 0xffff")
        self.e(f"s_cmp_eq_u32 s{SCR}, 0")
        self.e(f"s_cbranch_scc1 {self.lbl(tag)}")
        self.e(f"s_lshr_b32 s{SCR + 1}, s{HDR + 2}, 16")
        self.e(f"v_and_b32 %[vt], s{SCR}, %[tid]")
        self.e(f"v_cmp_eq_u32 vcc, s{SCR + 1}, %[vt]")
        self.e("s_and_b64 exec, exec, vcc")
        self.e(f"{self.lbl(tag)}:")

    def neg(self):
        f = self.sfx
        for j in range(self.NR):
            tag = f"ng{j}"
            self.jskip(j, tag)
            self.e(f"v_mul_{f} {self.re(j)}, -1.0, {self.re(j)}")
            self.e(f"v_mul_{f} {self.im(j)}, -1.0, {self.im(j)}")
            self.e(f"{self.lbl(tag)}:")

    def phase(self):
        f = self.sfx
        pr, pi = self.sreg(0), self.sreg(1)
        for j in range(self.NR):
            tag = f"ph{j}"
            self.jskip(j, tag)
            self.e(f"v_mul_{f} %[t0], {pi}, {self.im(j)}")
            self.e(f"v_mul_{f} %[t1], {pi}, {self.re(j)}")
            self.e(f"v_fma_{f} {self.re(j)}, {pr}, {self.re(j)}, -%[t0]")
            self.e(f"v_fma_{f} {self.im(j)}, {pr}, {self.im(j)}, %[t1]")
            self.e(f"{self.lbl(tag)}:")

    def mat(self, B):
        f = self.sfx
        if self.d:
            u = [self.sreg(i) for i in range(4)] + [self.sbuf(BUF2, i) for i in range(4)]
        else:
            u = [self.sreg(i) for i in range(8)]
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

    def swap(self, B):
        for j0, j1 in self.pairs(B):
            tag = f"sw{B}_{j0}"
            self.jskip(j0, tag)
            for a, b in ((self.re(j0), self.re(j1)), (self.im(j0), self.im(j1))):
                self.e(f"{self.mov} %[t0], {a}")
                self.e(f"{self.mov} {a}, {b}")
                self.e(f"{self.mov} {b}, %[t0]")
            self.e(f"{self.lbl(tag)}:")

    def rmat(self, B):
        f = self.sfx
        m0, m1, m2, m3 = (self.sreg(i) for i in range(4))
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

    def per_bit(self, kind, body):
        """dispatch on the register bit (header dword 0 = kind * 8 + bit), then body(B)."""
        for B in range(self.R):
            self.e(f"s_cmp_eq_u32 s{HDR}, {KIND[kind] * 8 + B}")
            self.e(f"s_cbranch_scc1 {self.lbl(f'{kind}{B}')}")
        self.e(f"s_branch {self.lbl('done')}")
        for B in range(self.R):
            self.e(f"{self.lbl(f'{kind}{B}')}:")
            body(B)
            self.e(f"s_branch {self.lbl('done')}")

    # --- the whole interpreter -----------------------------------------------
    def build(self):
        e = self.e
        e(f"s_mov_b64 {sp(NEXT)}, %[prog]")
        e(f"s_mov_b32 s{CNT}, %[nops]")
        e(f"s_cmp_eq_u32 s{CNT}, 0")
        e(f"s_cbranch_scc1 {self.lbl('end')}")
        e(f"{self.lbl('op')}:")
        e(f"s_load_dwordx16 s[{HDR}:{HDR + 15}], {sp(NEXT)}, 0x0")
        e("s_waitcnt lgkmcnt(0)")
        e(f"s_mov_b64 {sp(CUR)}, {sp(NEXT)}")
        e(f"s_add_u32 s{NEXT}, s{NEXT}, s{HDR + 3}")
        e(f"s_addc_u32 s{NEXT + 1}, s{NEXT + 1}, 0")
        e(f"s_sub_u32 s{CNT}, s{CNT}, 1")
        e(f"s_and_b64 {sp(SCR)}, %[base], {sp(HDR + 4)}")
        e(f"s_cmp_eq_u64 {sp(SCR)}, {sp(HDR + 6)}")
        e(f"s_cbranch_scc0 {self.lbl('next')}")
        e(f"s_mov_b64 {sp(EXEC)}, exec")
        for kind in ("SHEAR", "DIAGR", "NEG", "PHASE", "MAT", "SWAP", "RMAT"):
            e(f"s_cmp_lt_u32 s{HDR}, {8 * (KIND[kind] + 1)}")
            e(f"s_cbranch_scc1 {self.lbl('k' + kind)}")
        e(f"s_branch {self.lbl('next')}")

        e(f"{self.lbl('kSHEAR')}:")
        self.per_bit("SHEAR", self.shear)

        e(f"{self.lbl('kDIAGR')}:")
        self.diagr()
        e(f"s_branch {self.lbl('done')}")

        e(f"{self.lbl('kNEG')}:")
        self.lanes_on("nglanes")
        self.neg()
        e(f"s_branch {self.lbl('done')}")

        e(f"{self.lbl('kPHASE')}:")
        self.lanes_on("phlanes")
        self.phase()
        e(f"s_branch {self.lbl('done')}")

        e(f"{self.lbl('kMAT')}:")
        if self.d:  # u10, u11 from the second 32 bytes of coefficients
            e(f"s_load_dwordx8 s[{BUF2}:{BUF2 + 7}], {sp(CUR)}, 0x40")
        self.lanes_on("mtlanes")
        if self.d:
            e("s_waitcnt lgkmcnt(0)")
        self.per_bit("MAT", self.mat)

        e(f"{self.lbl('kSWAP')}:")
        self.lanes_on("swlanes")
        self.per_bit("SWAP", self.swap)

        e(f"{self.lbl('kRMAT')}:")
        self.per_bit("RMAT", self.rmat)

        e(f"{self.lbl('done')}:")
        e(f"s_mov_b64 exec, {sp(EXEC)}")
        e(f"{self.lbl('next')}:")
        e(f"s_cmp_lg_u32 s{CNT}, 0")
        e(f"s_cbranch_scc1 {self.lbl('op')}")
        e(f"{self.lbl('end')}:")
        return self.L

    def emit(self):
        body = self.build()
        out = []
        out.append(f"template <> struct AsmOps<{self.T}, {self.R}> {{")
        out.append(f"    __device__ __forceinline__ static void run(Regs<{self.T}, {self.R}> &a, const void *prog,")
        out.append("                                               uint32_t nops, uint64_t base, uint32_t tid) {")
        out.append(f"        {self.T} t0, t1, t2, t3;")
        out.append("        uint32_t vt;")
        out.append("        asm volatile(")
        for ln in body:
            pre = "" if ln.endswith(":") else "\\t"
            out.append(f'            "{pre}{ln}\\n"')
        outs = [f'[re{j}] "+v"(a.re[{j}])' for j in range(self.NR)] + \
               [f'[im{j}] "+v"(a.im[{j}])' for j in range(self.NR)] + \
               ['[t0] "=&v"(t0)', '[t1] "=&v"(t1)', '[t2] "=&v"(t2)', '[t3] "=&v"(t3)', '[vt] "=&v"(vt)']
        ins = ['[prog] "s"(prog)', '[nops] "s"(nops)', '[base] "s"(base)', '[tid] "v"(tid)']
        out.append("            : " + ",\n              ".join(outs))
        out.append("            : " + ", ".join(ins))
        clob = ", ".join(f'"s{i}"' for i in CLOBBER)
        out.append(f'            : {clob}, "vcc", "scc", "memory");')
        out.append("    }")
        out.append("};")
        return "\n".join(out)


def main():
    print("// Generated by gen_ops_asm.py -- the sweep kernel's op interpreter in gfx950")
    print("// assembly, one body per (amplitude type, register bits).  Do not edit.")
    print("template <typename T, int R> struct AsmOps;")
    for T in ("double", "float"):
        for R in range(1, 6):
            print(Gen(T, R).emit())
            print()


if __name__ == "__main__":
    sys.exit(main())
