#!/usr/bin/env python3
"""Emit the FHP collision rule as a Boolean circuit of 3-input lookup nodes.

The rule is the reference's: every state is grouped with the states that have
the same mass and the same exact momentum (src/collision.rs:37-69, classes in
(mass, px2, py2) key order, members ascending), and a collision replaces a
state by member ``k % |g|`` of its class, where ``k = (r * stride) >> 32`` is
the CPU table's column (src/collision.rs:86-89,115-119). This generator turns
that enumeration into straight-line bitplane code, so that one evaluation
collides 64 cells at once and agrees with the CPU table for every state and
every ``k``.

Inputs of the emitted function, one bit per cell:
  n[0..6]  the gathered state planes (six directions, then rest)
  K[0..B)  the binary digits of k (B = 5 for stride 30, 3 for stride 6)
Output: o[0..6], the collided state planes.

Construction (this generator's own; the reference's Metal emitter builds a
prefix tree instead):
  * k is decoded once per word into one-hot planes S2[m] = (k%2 == m),
    S3[m] = (k%3 == m), S5[m] = (k%5 == m), using that k < stride;
  * the state minterms needed by non-singleton classes are products of three
    two-literal pair products (bits 0-1, 2-3, 4-5) and the rest literal;
  * each class indicator i_c is the OR of its members' minterms;
  * output d is (n_d & keep) | OR over (size, m) of S_size[m] & U[d][size][m],
    where U is the OR of the indicators of the classes of that size whose m-th
    member has bit d set, and keep = ~OR_c i_c.
The netlist is then folded greedily into nodes of at most three inputs, each
carrying an 8-bit truth table indexed by (a << 2) | (b << 1) | c, which is the
operand order of v_bitop3_b32 on gfx950.
"""
from __future__ import annotations

import argparse
import itertools
import os
import sys

CX2 = [2, 1, -1, -2, -1, 1]   # src/hex.rs:29
CY2 = [0, 1, 1, 0, -1, -1]    # src/hex.rs:30
REST_BIT = 1 << 6


def invariants(s: int):
    """src/collision.rs:37-52."""
    mass = px = py = 0
    for d in range(6):
        if s & (1 << d):
            mass += 1
            px += CX2[d]
            py += CY2[d]
    if s & REST_BIT:
        mass += 1
    return (mass, px, py)


def classes(rest: bool):
    """src/collision.rs:62-69: sorted by key, members ascending."""
    groups = {}
    for s in range(128 if rest else 64):
        groups.setdefault(invariants(s), []).append(s)
    return [groups[k] for k in sorted(groups)]


def stride_of(rest: bool) -> int:
    from math import gcd
    st = 1
    for g in classes(rest):
        st = st * len(g) // gcd(st, len(g))
    return st


# ---------------------------------------------------------------------------
# A netlist of <=3-input truth-table nodes
# ---------------------------------------------------------------------------

class Net:
    def __init__(self):
        self.nodes = []        # (kind, payload); kind in {"in", "op"}
        self.hash = {}

    def inp(self, name):
        self.nodes.append(("in", name))
        return len(self.nodes) - 1

    def op(self, tt, *args):
        """A node computing truth table `tt` over 1..3 argument node ids."""
        args = list(args)
        # canonicalise: pad to 3 args by duplicating? keep arity; dedupe structure
        key = (tt, tuple(args))
        if key in self.hash:
            return self.hash[key]
        self.nodes.append(("op", (tt, args)))
        self.hash[key] = len(self.nodes) - 1
        return len(self.nodes) - 1

    # 2-input helpers with literal complements folded in
    def and2(self, a, b, na=False, nb=False):
        tt = 0
        for x, y in itertools.product((0, 1), repeat=2):
            if (x ^ na) and (y ^ nb):
                tt |= 1 << ((x << 1) | y)
        return self.op(tt, a, b)

    def or2(self, a, b):
        return self.op(0b1110, a, b)

    def orn(self, xs):
        """OR of a list, as a balanced tree of 2-input ORs (folded later)."""
        xs = list(xs)
        assert xs
        while len(xs) > 1:
            nxt = []
            for i in range(0, len(xs) - 1, 2):
                nxt.append(self.or2(xs[i], xs[i + 1]))
            if len(xs) % 2:
                nxt.append(xs[-1])
            xs = nxt
        return xs[0]

    def not1(self, a):
        return self.op(0b01, a)

    def andor(self, s, u, acc):
        """acc | (s & u)"""
        return self.op(self._tt3(lambda a, b, c: (a & b) | c), s, u, acc)

    @staticmethod
    def _tt3(f):
        tt = 0
        for a, b, c in itertools.product((0, 1), repeat=3):
            if f(a, b, c):
                tt |= 1 << ((a << 2) | (b << 1) | c)
        return tt


def eval_tt(tt, vals):
    """Truth table over len(vals) inputs, index built MSB-first."""
    idx = 0
    for v in vals:
        idx = (idx << 1) | v
    return (tt >> idx) & 1


def fold(net: Net, outputs):
    """Greedy merge of single-fanout (or 1-input) nodes into their consumers
    while the merged node keeps at most three distinct inputs."""
    # represent op nodes as (tt over arity, args)
    ops = {}
    for i, (kind, p) in enumerate(net.nodes):
        if kind == "op":
            ops[i] = (p[0], list(p[1]))

    def fanout():
        fo = {i: 0 for i in range(len(net.nodes))}
        for i, (tt, args) in ops.items():
            for a in set(args):
                fo[a] += 1
        for o in outputs:
            fo[o] += 1
        return fo

    def compose(tt_n, args_n, j, tt_m, args_m):
        """Inline node j (tt_m over args_m) into node (tt_n over args_n)."""
        new_args = []
        for a in args_n:
            if a == j:
                for b in args_m:
                    if b not in new_args:
                        new_args.append(b)
            elif a not in new_args:
                new_args.append(a)
        if len(new_args) > 3:
            return None
        tt = 0
        for bits in itertools.product((0, 1), repeat=len(new_args)):
            env = dict(zip(new_args, bits))
            mval = eval_tt(tt_m, [env[b] for b in args_m])
            vals = [mval if a == j else env[a] for a in args_n]
            if eval_tt(tt_n, vals):
                idx = 0
                for b in bits:
                    idx = (idx << 1) | b
                tt |= 1 << idx
        return tt, new_args

    changed = True
    while changed:
        changed = False
        fo = fanout()
        for i in sorted(ops):
            tt_n, args_n = ops[i]
            for j in list(dict.fromkeys(args_n)):
                if j not in ops:
                    continue
                tt_m, args_m = ops[j]
                if fo[j] != 1 and len(args_m) != 1:
                    continue
                r = compose(tt_n, args_n, j, tt_m, args_m)
                if r is None:
                    continue
                ops[i] = r
                tt_n, args_n = r
                changed = True
            if changed:
                break
    # drop dead nodes
    live = set()
    stack = list(outputs)
    while stack:
        x = stack.pop()
        if x in live:
            continue
        live.add(x)
        if x in ops:
            stack.extend(ops[x][1])
    return {i: ops[i] for i in ops if i in live}


def _decompose2(f, care, nin):
    """Find f = g(h(x_A), x_B) with |A| = 3, |B| = nin - 3 on the care set.
    Returns (A, B, h_tt, g_tt) or None. Truth tables index MSB-first."""
    from itertools import combinations
    for A in combinations(range(nin), 3):
        B = [i for i in range(nin) if i not in A]
        cols = {}
        for x in range(1 << nin):
            if not care(x):
                continue
            a = tuple((x >> i) & 1 for i in A)
            b = tuple((x >> i) & 1 for i in B)
            cols.setdefault(a, {})[b] = f(x)
        keys = list(cols)
        for mask in range(1 << len(keys)):
            grp = [{}, {}]
            ok = True
            for bi, key in enumerate(keys):
                g = grp[(mask >> bi) & 1]
                for b, v in cols[key].items():
                    if g.get(b, v) != v:
                        ok = False
                        break
                    g[b] = v
                if not ok:
                    break
            if not ok:
                continue
            h_tt = 0
            for bi, key in enumerate(keys):
                if (mask >> bi) & 1:
                    idx = 0
                    for v in key:
                        idx = (idx << 1) | v
                    h_tt |= 1 << idx
            g_tt = 0
            for hv in (0, 1):
                for b, v in grp[hv].items():
                    if v:
                        idx = hv
                        for bit in b:
                            idx = (idx << 1) | bit
                        g_tt |= 1 << idx
            return list(A), B, h_tt, g_tt
    return None


def build_rule(rest: bool, grouped: bool = True):
    net = Net()
    n = [net.inp(f"n[{d}]") for d in range(7)]
    st = stride_of(rest)
    nk = (st - 1).bit_length()
    K = [net.inp(f"K[{b}]") for b in range(nk)]
    nbits = 7 if rest else 6
    cls = [g for g in classes(rest) if len(g) > 1]
    sizes = sorted({len(g) for g in cls})
    assert all(st % s == 0 for s in sizes)
    assert set(sizes) <= {2, 3, 5}

    # --- one-hot residues of k, digit by digit: R[m] = [k_partial % q == m].
    # The first three digits are one node per residue; each further digit
    # 2^i shifts the residue by 2^i mod q, a 3-input select per residue.
    def residues(q):
        lo = min(3, nk)
        R = []
        for m in range(q):
            tt = 0
            for y in range(1 << lo):           # y = digits 0..lo-1
                if y % q == m:
                    idx = 0
                    for i in reversed(range(lo)):  # args: K[lo-1], ..., K[0]
                        idx = (idx << 1) | ((y >> i) & 1)
                    tt |= 1 << idx
            R.append(net.op(tt, *[K[i] for i in reversed(range(lo))]) if tt else None)
        for i in range(lo, nk):
            sh = (1 << i) % q
            R2 = []
            for m in range(q):
                a, b = R[(m - sh) % q], R[m]       # digit set: came from m - sh
                if a is None and b is None:
                    R2.append(None)
                elif a is None:
                    R2.append(net.and2(K[i], b, True, False))
                elif b is None:
                    R2.append(net.and2(K[i], a))
                else:
                    R2.append(net.op(Net._tt3(lambda d, x, y: x if d else y), K[i], a, b))
            R = R2
        return R

    res = {2: None, 3: residues(3) if 3 in sizes else None, 5: residues(5) if 5 in sizes else None}

    # sel(size, M) -> (node, negated) for "k % size in M", or "one"/None.
    pair_cache = {}

    def sel(size, M):
        M = sorted(set(M))
        if not M:
            return None
        if len(M) == size:
            return "one"
        if size == 2:
            return (K[0], M == [0])                # k even <=> digit 0 clear
        R = res[size]
        comp = [m for m in range(size) if m not in M]
        if len(M) == 1:
            return (R[M[0]], False)
        if len(comp) == 1:
            return (R[comp[0]], True)
        use, neg = (M, False) if len(M) <= len(comp) else (comp, True)
        key = (size, tuple(use))
        if key not in pair_cache:
            pair_cache[key] = net.orn([R[m] for m in use])
        return (pair_cache[key], neg)

    # --- state minterms: pair products over bits (0,1), (2,3) and a triple
    # product over bits (4,5,6), then one 3-input AND per state.
    def lit_tt(vals):
        tt = 0
        for idx in range(1 << len(vals)):
            bits = [(idx >> (len(vals) - 1 - i)) & 1 for i in range(len(vals))]
            if bits == list(vals):
                tt |= 1 << idx
        return tt

    prod = {}

    def product(bits):  # bits: tuple of (plane, value)
        key = tuple(bits)
        if key not in prod:
            prod[key] = net.op(lit_tt([v for _, v in bits]), *[n[p] for p, _ in bits])
        return prod[key]

    needed = sorted({s for g in cls for s in g})
    mins = {}
    for s in needed:
        a = product(((0, s & 1), (1, (s >> 1) & 1)))
        b = product(((2, (s >> 2) & 1), (3, (s >> 3) & 1)))
        if rest:
            c = product(((4, (s >> 4) & 1), (5, (s >> 5) & 1), (6, (s >> 6) & 1)))
        else:
            c = product(((4, (s >> 4) & 1), (5, (s >> 5) & 1)))
        mins[s] = net.op(Net._tt3(lambda x, y, z: x & y & z), a, b, c)

    ind = [net.orn([mins[s] for s in g]) for g in cls]
    anyc = net.orn(ind)

    def andor(sv, u, acc):
        node, neg = sv
        return net.op(Net._tt3(lambda a, b, c: ((a ^ neg) & b) | c), node, u, acc)

    outs = []
    for d in range(7):
        if d >= nbits:
            outs.append(n[d])
            continue
        acc = net.and2(n[d], anyc, False, True)      # n_d & keep
        groups = {}
        for ci, g in enumerate(cls):
            M = tuple(m for m in range(len(g)) if (g[m] >> d) & 1)
            if M:
                groups.setdefault((len(g), M), []).append(ci)
        for (size, M), cis in sorted(groups.items()):
            sv = sel(size, M)
            parts = [ind[ci] for ci in cis] if grouped else None
            if grouped:
                u = net.orn(parts)
                acc = net.or2(acc, u) if sv == "one" else andor(sv, u, acc)
            else:
                for ci in cis:
                    acc = net.or2(acc, ind[ci]) if sv == "one" else andor(sv, ind[ci], acc)
        outs.append(acc)
    return net, outs, st, nk


def emit(rest: bool, fname: str):
    best = None
    for grouped in (True, False):
        net_, outs_, st, nk = build_rule(rest, grouped)
        ops_ = fold(net_, outs_)
        if best is None or len(ops_) < len(best[2]):
            best = (net_, outs_, ops_)
    net, outs, ops = best
    order = sorted(ops)  # node ids are topological by construction
    names = {}
    for i, (kind, p) in enumerate(net.nodes):
        if kind == "in":
            names[i] = p
    lines = []
    for i in order:
        tt, args = ops[i]
        a = [names[x] for x in args]
        if len(a) == 1:
            a3, tt3 = [a[0], a[0], a[0]], 0
            for x in (0, 1):
                if eval_tt(tt, [x]):
                    tt3 |= 1 << ((x << 2) | (x << 1) | x)
            # fill the unreachable indices freely: use the 1-input table on bit a
            tt3 = 0
            for idx in range(8):
                if eval_tt(tt, [(idx >> 2) & 1]):
                    tt3 |= 1 << idx
        elif len(a) == 2:
            a3 = [a[0], a[1], a[1]]
            tt3 = 0
            for idx in range(8):
                if eval_tt(tt, [(idx >> 2) & 1, (idx >> 1) & 1]):
                    tt3 |= 1 << idx
        else:
            a3, tt3 = a, tt
        names[i] = f"t{i}"
        lines.append(f"  const W t{i} = lg_op3<0x{tt3:02x}>({a3[0]}, {a3[1]}, {a3[2]});")
    for d in range(7):
        lines.append(f"  o[{d}] = {names[outs[d]]};")
    body = "\n".join(lines)
    sig = (f"template <typename W>\nLG_HD inline void {fname}(const W n[7], const W K[{nk}], W o[7])")
    return (f"// {len(order)} three-input nodes; stride {st}; k has {nk} binary digits.\n"
            f"{sig} {{\n{body}\n}}\n"), len(order), st, nk


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    parts = []
    info = []
    for rest, fname in ((True, "lg_collide_rest"), (False, "lg_collide_norest")):
        src, nops, st, nk = emit(rest, fname)
        parts.append(src)
        info.append((fname, nops, st, nk))
    hdr = ["// GENERATED by lattice-gas_amd/gen_collide.py -- do not edit.",
           "// The FHP collision rule of src/collision.rs:37-119 as bitplane code.",
           "// Requires LG_HD and lg_op3<table>(a, b, c) (32- or 64-bit words) from lgca_device.h.",
           "#pragma once", ""]
    for fname, nops, st, nk in info:
        hdr.append(f"#define {fname.upper()}_NODES {nops}")
        hdr.append(f"#define {fname.upper()}_STRIDE {st}")
        hdr.append(f"#define {fname.upper()}_KBITS {nk}")
    hdr.append("")
    text = "\n".join(hdr) + "\n" + "\n".join(parts)
    tmp = args.out + ".tmp"
    with open(tmp, "w") as f:
        f.write(text)
    os.replace(tmp, args.out)
    for fname, nops, st, nk in info:
        print(f"{fname}: {nops} nodes, stride {st}, {nk} k-digits", file=sys.stderr)


if __name__ == "__main__":
    main()
