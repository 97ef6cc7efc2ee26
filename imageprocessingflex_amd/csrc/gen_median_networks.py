"""Generates median_networks.hpp: min/max selection networks for the median
kernels, built from Batcher's odd-even merge networks and pruned.

  python gen_median_networks.py        # rewrites median_networks.hpp

Emitted functions (K = key type, v = registers, statically indexed):
  sort3(a, b, c)                 -> a <= b <= c
  sort5(v[5])                    -> v ascending
  med25_cols(v[25])              -> median of 25 values given as 5 columns
                                    v[5*k .. 5*k+4], each column ascending

Every network is checked here by the 0-1 principle: a comparator network
selects rank r correctly for all inputs iff it does so for every 0-1 input
that satisfies its precondition (sorted columns). The checks are exhaustive.
"""
from __future__ import annotations

import itertools
from pathlib import Path

HERE = Path(__file__).resolve().parent


def oddeven_merge(lo, hi, r):
    step = r * 2
    if step < hi - lo:
        yield from oddeven_merge(lo, hi, step)
        yield from oddeven_merge(lo + r, hi, step)
        yield from [(i, i + r) for i in range(lo + r, hi - r, step)]
    else:
        yield (lo, lo + r)


def oddeven_sort(lo, hi):
    if hi - lo >= 1:
        mid = lo + (hi - lo) // 2
        yield from oddeven_sort(lo, mid)
        yield from oddeven_sort(mid + 1, hi)
        yield from oddeven_merge(lo, hi, 1)


def merge_block(lo, n):
    """Odd-even merge of the two sorted halves of wires [lo, lo + n)."""
    return [(lo + i, lo + j) for i, j in oddeven_merge(0, n - 1, 1)]


def lower(net, finite, n_wires):
    """Turns a network over padded wires (+inf where not finite) into a list of
    ('cx', i, j) and ('mv', dst, src) ops over finite wires only."""
    inf = [not finite[w] for w in range(n_wires)]
    ops = []
    for i, j in net:
        if inf[i] and inf[j]:
            continue
        if inf[j]:  # min(x, inf) = x stays on i
            continue
        if inf[i]:  # the finite value moves down to i
            ops.append(("mv", i, j))
            inf[i], inf[j] = False, True
            continue
        ops.append(("cx", i, j))
    return ops, [not x for x in inf]


def prune(ops, outputs):
    """Backward liveness: keeps only ops that feed the given output wires."""
    live = set(outputs)
    kept = []
    for op in reversed(ops):
        k, i, j = op
        if k == "mv":
            if i in live:
                live.discard(i)
                live.add(j)
                kept.append(op)
            continue
        if i in live or j in live:
            kept.append((k, i, j, i in live, j in live))
            live.add(i)
            live.add(j)
    kept.reverse()
    return kept, live


def run(ops, vals):
    v = list(vals)
    for op in ops:
        if op[0] == "mv":
            v[op[1]] = v[op[2]]
        else:
            _, i, j, *_ = op
            a, b = v[i], v[j]
            v[i], v[j] = min(a, b), max(a, b)
    return v


def build_sort(n):
    n2 = 1
    while n2 < n:
        n2 *= 2
    ops, _ = lower(list(oddeven_sort(0, n2 - 1)), [w < n for w in range(n2)], n2)
    ops, _ = prune(ops, range(n))
    for bits in itertools.product((0, 1), repeat=n):  # 0-1 principle
        assert run(ops, list(bits) + [9] * (n2 - n))[:n] == sorted(bits)
    return ops


def build_med25_cols():
    # column k on wires [8k, 8k + 5); wires 8k+5..8k+7 and 40..63 are +inf
    n2 = 64
    finite = [False] * n2
    for k in range(5):
        for i in range(5):
            finite[8 * k + i] = True
    net = merge_block(0, 16) + merge_block(16, 16) + merge_block(0, 32) + merge_block(0, 64)
    ops, _ = lower(net, finite, n2)
    ops, _ = prune(ops, [12])
    # exhaustive check over sorted 0-1 columns (6^5 cases)
    for ones in itertools.product(range(6), repeat=5):
        vals = [9] * n2
        allv = []
        for k, o in enumerate(ones):
            col = [0] * (5 - o) + [1] * o
            allv += col
            for i in range(5):
                vals[8 * k + i] = col[i]
        assert run(ops, vals)[12] == sorted(allv)[12], ones
    # wire map: padded wire 8k+i <- input index 5k+i
    return ops, {8 * k + i: 5 * k + i for k in range(5) for i in range(5)}


def emit(ops, name, arr, nin, wiremap, out_wire, ret=True):
    lines = []
    wires = sorted({w for op in ops for w in op[1:3]} | set(wiremap) | {out_wire})
    lines.append(f"template <class K>\n__device__ __forceinline__ {'K' if ret else 'void'} {name}({'const ' if ret else ''}K (&{arr})[{nin}]) {{")
    if ret:
        for w in wires:
            src = wiremap.get(w)
            if src is not None:
                lines.append(f"    K w{w} = {arr}[{src}];")
            else:
                lines.append(f"    K w{w};")
    for op in ops:
        if op[0] == "mv":
            lines.append(f"    w{op[1]} = w{op[2]};")
            continue
        _, i, j, need_i, need_j = op
        lines.append(f"    {{ const K a = w{i}, b = w{j};"
                     + (f" w{i} = kmin(a, b);" if need_i else "")
                     + (f" w{j} = kmax(a, b);" if need_j else "") + " }")
    lines.append(f"    return w{out_wire};\n}}\n")
    return "\n".join(lines)


def emit_sort(ops, n):
    lines = [f"template <class K>\n__device__ __forceinline__ void sort{n}(K (&v)[{n}]) {{"]
    for op in ops:
        if op[0] == "mv":
            lines.append(f"    v[{op[1]}] = v[{op[2]}];")
            continue
        _, i, j, need_i, need_j = op
        lines.append(f"    {{ const K a = v[{i}], b = v[{j}]; v[{i}] = kmin(a, b); v[{j}] = kmax(a, b); }}")
    lines.append("}\n")
    return "\n".join(lines)


def main():
    s5 = build_sort(5)
    m25, wm = build_med25_cols()
    n_cx = sum(1 for op in m25 if op[0] == "cx")
    n_ops = sum((op[3] + op[4]) for op in m25 if op[0] == "cx")
    out = [
        "// GENERATED by gen_median_networks.py -- do not edit.",
        "// Batcher odd-even merge networks, padded with +inf wires, lowered and",
        "// pruned to the wires that reach the output; verified exhaustively by",
        "// the 0-1 principle in the generator.",
        f"// sort5: {len(s5)} comparators; med25_cols: {n_cx} comparators, {n_ops} min/max ops.",
        "#pragma once",
        "",
        emit_sort(s5, 5),
        emit(m25, "med25_cols", "v", 25, wm, 12),
    ]
    (HERE / "median_networks.hpp").write_text("\n".join(out))
    print(f"sort5 {len(s5)} comparators; med25_cols {n_cx} comparators / {n_ops} ops")


if __name__ == "__main__":
    main()
