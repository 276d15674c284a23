"""CPU restatement of the ligate tool's phase alignment -- test infrastructure
only (the checker of glimpse_amd's gh_ligate), never part of the product path.

Follows ligate/src/ligater/ligater_algorithm.cpp: update_distances (:108-127)
per overlap record and sample, the swap decision and phase quality of
scan_overlap (:196-213), phase_update (:90-106), and the record order of the
main loop (:339-444): chunk f's records before chunk f + 1 starts, then the
synced walk over the overlap (records in both: the first L_isec / 2 from f,
the rest from f + 1; records of one chunk from it), then chunk f + 1.  The
swap state carries over between chunks without an overlap.

Where the reference's synced reader would use the previous overlap's swap
state for records present in chunk f + 1 only that come before the first
record common to both (a state the first common record then replaces), each
chunk's own swap state is used here, as the GPU path does.

Pinned in tests/test_ligate_oracle.py: chunks cut from one phased truth with
known per-sample phase flips must be un-flipped exactly, so that the ligated
output equals the truth; the phase quality against its closed form.
"""
import math

import numpy as np

VECTOR_END = -2147483647


def _missing(v):
    return (v >> 1) == 0


def _allele(v):
    return (v >> 1) - 1


def update_distances(gta, gtb, swap_a, nmatch, nmism):
    """One overlap record: gta, gtb [N][2] int."""
    for i in range(len(swap_a)):
        a0, a1 = int(gta[i][0]), int(gta[i][1])
        b0, b1 = int(gtb[i][0]), int(gtb[i][1])
        if a1 == VECTOR_END or b1 == VECTOR_END:
            continue
        if _missing(a0) or _missing(a1) or _missing(b0) or _missing(b1):
            continue
        if not (a1 & 1) or not (b1 & 1):
            continue
        if _allele(a0) == _allele(a1) or _allele(b0) == _allele(b1):
            continue
        if _allele(a0) == _allele(b0) and _allele(a1) == _allele(b1):
            if swap_a[i]:
                nmism[i] += 1
            else:
                nmatch[i] += 1
        if _allele(a0) == _allele(b1) and _allele(a1) == _allele(b0):
            if swap_a[i]:
                nmatch[i] += 1
            else:
                nmism[i] += 1


def phase_quality(nmatch, nmism):
    f = np.float32(99)
    if nmatch and nmism:
        f0 = np.float32(np.float32(nmatch) / np.float32(nmatch + nmism))
        # 0.7 + (float) f0 ln f0 is a double; the float (1 - f0) ln(1 - f0) is added to it
        t0 = float(np.float32(f0 * np.log(f0)))
        t1 = float(np.float32(np.float32(1 - f0) * np.log(np.float32(1 - f0))))
        f = np.float32(99 * ((0.7 + t0) + t1) / 0.7)
    return f


def phase_update(gt, swap):
    """gt [N][2] (one record), in place."""
    for i in range(len(swap)):
        if not swap[i]:
            continue
        g0, g1 = int(gt[i][0]), int(gt[i][1])
        if _missing(g0) or g1 == VECTOR_END:
            continue
        if not (g1 & 1):
            continue
        gt[i][0] = ((1 if _allele(g1) == 1 else 0) + 1) << 1 | 1
        gt[i][1] = ((1 if _allele(g0) == 1 else 0) + 1) << 1 | 1


def ligate(chunks, n_samples):
    """Returns (ligated GT rows [n_out][N][2], order [(chunk, record)], swap per chunk)."""
    N = n_samples
    nc = len(chunks)
    swaps = [np.zeros(N, bool)]
    plans = []
    for f in range(nc - 2):
        A, C = chunks[f], chunks[f + 2]
        if len(A["pos"]) and len(C["pos"]) and C["contig"][0] == A["contig"][-1] and C["pos"][0] <= A["pos"][-1]:
            raise ValueError("Three files overlapping")   # ligater_algorithm.cpp:386
    for f in range(nc - 1):
        A, B = chunks[f], chunks[f + 1]
        plan = None
        if len(A["pos"]) and len(B["pos"]) and A["contig"][0] == B["contig"][0]:
            ctg, start = B["contig"][0], B["pos"][0]
            idx = [r for r in range(len(A["pos"])) if A["contig"][r] == ctg and A["pos"][r] >= start]
            if idx:
                ia = idx[0]
                seek = A["pos"][ia]
                ib = 0
                while ib < len(B["pos"]) and B["contig"][ib] == ctg and B["pos"][ib] < seek:
                    ib += 1
                recs = []
                used = set()
                ra = ia
                while ra < len(A["pos"]):
                    pos = A["pos"][ra]
                    ea = ra
                    while ea < len(A["pos"]) and A["pos"][ea] == pos:
                        ea += 1
                    while ib < len(B["pos"]) and B["pos"][ib] < pos:
                        recs.append((None, ib))
                        ib += 1
                    eb = ib
                    while eb < len(B["pos"]) and B["pos"][eb] == pos:
                        eb += 1
                    for r in range(ra, ea):
                        hit = None
                        for q in range(ib, eb):
                            if q not in used and B["key"][q] == A["key"][r]:
                                hit = q
                                break
                        if hit is not None:
                            used.add(hit)
                        recs.append((r, hit))
                    for q in range(ib, eb):
                        if q not in used:
                            recs.append((None, q))
                    ib = eb
                    ra = ea
                plan = {"a_from": ia, "b_next": ib, "recs": recs}
        plans.append(plan)
        sa = swaps[f]
        if plan is None or not any(a is not None and b is not None for a, b in plan["recs"]):
            swaps.append(sa.copy())
            continue
        nmatch, nmism = [0] * N, [0] * N
        n_scan = 0
        for a, b in plan["recs"]:
            if a is None or b is None:
                continue
            if A.get("biallelic") is not None and not A["biallelic"][a]:
                continue
            update_distances(A["gt"][a], B["gt"][b], sa, nmatch, nmism)
            n_scan += 1
        if n_scan == 0:
            raise ValueError("Overlap is empty")
        plan["n_scan"] = n_scan
        plan["nmatch"], plan["nmism"] = nmatch, nmism
        swaps.append(np.array([nmatch[i] < nmism[i] for i in range(N)]))
    order = []
    start = 0
    for f in range(nc):
        plan = plans[f] if f < nc - 1 else None
        end = plan["a_from"] if plan else len(chunks[f]["pos"])
        order += [(f, r) for r in range(start, end)]
        if plan is None:
            start = 0
            continue
        half = plan.get("n_scan", 0) // 2
        k = 0
        for a, b in plan["recs"]:
            if a is not None and b is not None:
                order.append((f, a) if k < half else (f + 1, b))
                k += 1
            elif a is not None:
                order.append((f, a))
            else:
                order.append((f + 1, b))
        start = plan["b_next"]
    rows = []
    for c, r in order:
        g = np.array(chunks[c]["gt"][r], np.int64)
        phase_update(g, swaps[c])
        rows.append(g)
    out = np.array(rows, np.int32).reshape(len(order), N, 2)
    return out, order, swaps
