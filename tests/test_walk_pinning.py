"""Pins the oracle's cell walk (po_walk_query) to the reference's pruning rules.

PuckIndex::search_nearest_filter_points (puck/puck/puck_index.cpp:506-557) and
compute_quantized_distance (:346-390) decide, cell by cell, whether to stop a
coarse list (break), skip a cell (continue) or scan it, from a pivot that
follows the filter heap's top.  Three kinds of check:

* Hand-traced known answers.  Tiny indexes whose every walk input (coarse and
  fine distances, stationary values, offsets, codes, |q|^2, radius_rate) is a
  small dyadic number, so the float arithmetic is exact and each case was
  traced by hand from the reference lines (the traces are in the comments).
  Each case expects a filter set and walk counters.
* An independent trace checker.  po_walk_trace logs every considered cell;
  the checker recomputes every decision from the rules with its own
  arithmetic (numpy float32/float64) and its own top (the ftk-th smallest of
  all filter distances accepted so far, not a heap), on random indexes with
  radius_rate > 1 and filter_topk much smaller than N, so the pivot moves.
* Planted mistakes.  Each PO_MUT_* switch of puck_oracle.c alters one rule
  (pivot without /2, without /radius_rate or without -|q|^2, '>' for '>=' in
  the break, continue or point tests, max_stationary_dist not lowered or also
  lowered after a continue, its start value without the first fine distance,
  min_dist without min_dist_offset).  Each variant is compiled on the CPU and
  must fail the known answers or the checker.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

S = float(O.SENTINEL)
F32 = np.float32


# ---------------------------------------------------------------------------
# Tiny indexes with every walk input chosen.

def tiny_index(cd, fd, st, cells, L=None):
    """An index with dim = 2, nsq = 2 (sub-vectors of one float) and codebook
    rows cb[m][k] = k, searched with q = (0, 0): the lookup table entry of
    code k is k^2, so a point with codes (a, b) has LUT sum a^2 + b^2.
    cd [C] coarse distances (given as |c|^2/2 = 0 and <c,q> = -cd), fd [F]
    fine distances (<f,q> = -fd), st [C][F] stationary values, cells {(c, k):
    [(offset, (a, b)), ...]}.  Returns (index, walk arguments)."""
    C, F = len(cd), len(fd)
    recs, assign = [], []
    for (c, k), pts in sorted(cells.items()):
        for off, codes in pts:
            r = np.zeros(6, np.uint8)
            r[:4] = np.frombuffer(np.float32(off).tobytes(), np.uint8)
            r[4:] = codes
            recs.append(r)
            assign.append(c * F + k)
    records = np.stack(recs)
    raw = np.zeros((len(recs), 2), np.float32)
    cb = np.repeat(np.arange(256, dtype=np.float32)[None, :], 2, axis=0).reshape(2, 256, 1)
    ix = O.OracleIndex(np.zeros((C, 2), F32), np.zeros((F, 2), F32), cb, np.asarray(assign, np.uint32), records,
                       raw, False, L=L)
    ix.set_derived(np.zeros(C, F32), np.asarray(st, F32).reshape(-1))
    return ix, dict(q=np.zeros(2, F32), coarse_ip=-np.asarray(cd, F32), fine_ip=-np.asarray(fd, F32))


# Each case: walk inputs, then the expected filter set (sorted values) and
# counters (cells considered = break-tested, cells scanned, points whose LUT
# sum was taken, heap updates).
KATS = {
    # Pivot = (top - |q|^2) / radius_rate / 2 (puck_index.cpp:517, :544).
    # ftk 2, |q|^2 = 4, rr 1; cd = [1, 2]; fd = [0, 1, 3]; st c0 = [0, 1, 5],
    # c1 = [0.5, 0, 1]; min_dist_offset = [0, 0].  P0 = (S - 4)/2 = S/2.
    # l0 (cd 1, min_dist 1), max_st = P0:
    #   idx0 k0: 1 < P0, st 0 < P0: scan, cell_dist 1: off 0 -> 2 (heap {2}),
    #            off 10 -> 12 (heap {2, 12}, top 12) -> P = (12-4)/2 = 4;
    #            max_st = min(P0, 4 - 1 - 0) = 3
    #   idx1 k1: 1 + 1 = 2 < 4, st 1 < 3: scan, cell_dist 3: off 1 -> 7,
    #            top 7 -> P = 1.5; max_st = min(3, 1.5 - 1 - 1) = -0.5
    #   idx2 k2: 3 + 1 = 4 >= 1.5: break
    # l1 (cd 2, min_dist 2): idx0: 0 + 2 >= 1.5: break
    # -> {2, 7}; 4 considered, 2 scanned, 3 offered, 3 updates.  The points of
    # c1 (filter distances 2 and 3) are pruned: brute force would give {2, 2}.
    "pivot_half_and_query_norm": dict(
        cd=[1, 2], fd=[0, 1, 3], st=[[0, 1, 5], [0.5, 0, 1]], qn=4.0, rr=1.0, ftk=2, scc=2,
        cells={(0, 0): [(0, (0, 0)), (10, (0, 0))], (0, 1): [(1, (0, 0))], (0, 2): [(0, (0, 0))],
               (1, 0): [(-3, (0, 0))], (1, 1): [(-3, (0, 0))], (1, 2): [(0, (0, 0))]},
        filt=[2, 7], stats=(4, 2, 3, 3)),
    # radius_rate 2 and min_dist_offset (:522).  ftk 1, |q|^2 = 0; cd = [0, 1];
    # fd = [0, 0.5]; st c0 = [1, 1.5], c1 = [-0.5, 4]; mdo = [1, -0.5].
    # P0 = S/4.  l0 (cd 0, min_dist 1), max_st = S/4:
    #   idx0 k0: scan, cell_dist 1: off 2 -> 4 (top 4) -> P = 4/2/2 = 1;
    #            max_st = min(S/4, 1 - 0 - 0) = 1
    #   idx1 k1: 0.5 + 1 = 1.5 >= 1: break
    # l1 (cd 1, min_dist 0.5), max_st = 1 - 1 - 0 = 0:
    #   idx0 k0: 0.5 < 1, st -0.5 < 0: scan, cell_dist 0.5: off 0.5 -> 1.5
    #            (top 1.5) -> P = 0.375; max_st = min(0, 0.375 - 1 - 0)
    #   idx1 k1: 0.5 + 0.5 = 1 >= 0.375: break
    # -> {1.5}; 4 considered, 2 scanned, 2 offered, 2 updates.
    "radius_rate_and_min_dist_offset": dict(
        cd=[0, 1], fd=[0, 0.5], st=[[1, 1.5], [-0.5, 4]], qn=0.0, rr=2.0, ftk=1, scc=2,
        cells={(0, 0): [(2, (0, 0))], (0, 1): [(-3.5, (0, 0))], (1, 0): [(0.5, (0, 0))], (1, 1): [(0, (0, 0))]},
        filt=[1.5], stats=(4, 2, 2, 2)),
    # Ties at the break and continue tests, and max_stationary_dist lowered
    # only after a scanned cell (:534-547).  ftk 1, |q|^2 = 0, rr 1; one list,
    # cd 0; fd = [0, 1, 1.5, 1.75]; st = [0, 2, 1, 0]; mdo 0.  P0 = S/2.
    #   idx0 k0: scan, cell_dist 0: off 4 -> 4 (top 4) -> P = 2; max_st = 2
    #   idx1 k1: 1 < 2; st 2 >= 2: continue (max_st stays 2)
    #   idx2 k2: 1.5 < 2; st 1 < 2: scan, cell_dist 2.5: off -1.5 -> 3.5
    #            (top 3.5) -> P = 1.75; max_st = min(2, 1.75 - 0 - 1.5) = 0.25
    #   idx3 k3: 1.75 + 0 >= 1.75: break
    # -> {3.5}; 4 considered, 2 scanned, 2 offered, 2 updates.
    "ties_at_break_and_continue": dict(
        cd=[0], fd=[0, 1, 1.5, 1.75], st=[[0, 2, 1, 0]], qn=0.0, rr=1.0, ftk=1, scc=1,
        cells={(0, 0): [(4, (0, 0))], (0, 1): [(-3, (0, 0))], (0, 2): [(-1.5, (0, 0))], (0, 3): [(-3, (0, 0))]},
        filt=[3.5], stats=(4, 2, 2, 2)),
    # max_stationary_dist starts from pivot - cd - fine_distance[0] (:523-524),
    # the LUT sum, and the point test t0 >= top (:362-364).  ftk 1, |q|^2 0,
    # rr 1; cd = [0, 0.5]; fd = [-1, 0.5]; st c0 = [1, 10], c1 = [2, 5];
    # mdo = [1, 2].  P0 = S/2.
    # l0 (cd 0, min_dist 1):
    #   idx0 k0: scan, cell_dist 0: off 3, codes (1, 0) -> 3 + 1 = 4 (top 4);
    #            off 4 -> t0 = 4 >= 4: end of cell -> P = 2;
    #            max_st = min(S/2, 2 - 0 + 1) = 3
    #   idx1 k1: 0.5 + 1 = 1.5 < 2; st 10 >= 3: continue
    # l1 (cd 0.5, min_dist 2.5), max_st = 2 - 0.5 + 1 = 2.5:
    #   idx0 k0: -1 + 2.5 = 1.5 < 2; st 2 < 2.5: scan, cell_dist 1.5:
    #            off -0.5 -> 2.5 (top 2.5) -> P = 1.25; max_st = 1.75
    #   idx1 k1: 0.5 + 2.5 = 3 >= 1.25: break
    # -> {2.5}; 4 considered, 2 scanned, 2 offered, 2 updates.
    "max_st_start_and_point_break": dict(
        cd=[0, 0.5], fd=[-1, 0.5], st=[[1, 10], [2, 5]], qn=0.0, rr=1.0, ftk=1, scc=2,
        cells={(0, 0): [(3, (1, 0)), (4, (0, 0))], (0, 1): [(0, (0, 0))], (1, 0): [(-0.5, (0, 0))],
               (1, 1): [(0, (0, 0))]},
        filt=[2.5], stats=(4, 2, 2, 2)),
    # A heap that never fills: the pivot stays at the sentinel's, every cell is
    # scanned (the empty cell k1 too: its stationary sqrt(FLT_MAX) is below
    # max_st), and the heap is reordered ascending (:552-554).  ftk 8; one
    # list, cd 0; fd = [0, 1, 2]; st = [0, (empty), 1]; P = S/2 throughout.
    #   idx0 k0: scan, cell_dist 0: off 2 -> 2, off 3 -> 3
    #   idx1 k1: 1 < S/2; st sqrt(FLT_MAX) < S/2: scan, no points
    #   idx2 k2: scan, cell_dist 0 + 1 + 2 = 3: off 0 -> 6, off 3 -> 9
    # -> {2, 3, 6, 9}; 3 considered, 3 scanned, 4 offered, 4 updates.
    "heap_not_full": dict(
        cd=[0], fd=[0, 1, 2], st=[[0, 0, 1]], qn=0.0, rr=1.0, ftk=8, scc=1,
        cells={(0, 0): [(2, (0, 0)), (3, (0, 0))], (0, 2): [(0, (0, 0)), (3, (0, 0))]},
        filt=[2, 3, 6, 9], stats=(3, 3, 4, 4)),
}


def run_kat(name, L=None):
    k = KATS[name]
    ix, w = tiny_index(k["cd"], k["fd"], k["st"], k["cells"], L=L)
    steps, fv, ft, fn, st = ix.walk_trace(w["q"], w["coarse_ip"], w["fine_ip"], k["qn"], k["scc"], k["ftk"],
                                          k["rr"])
    return ix, w, steps, fv, ft, fn, st


def kat_result(name, L=None):
    _, _, _, fv, _, fn, st = run_kat(name, L)
    return sorted(fv[:fn].tolist()), (st["cells_considered"], st["cells_scanned"], st["points_offered"],
                                      st["heap_updates"])


@pytest.mark.parametrize("name", sorted(KATS))
def test_walk_known_answers(name):
    filt, stats = kat_result(name)
    assert filt == [float(v) for v in KATS[name]["filt"]], name
    assert stats == KATS[name]["stats"], name


def test_heap_not_full_is_reordered_ascending():
    _, _, _, fv, ft, fn, _ = run_kat("heap_not_full")
    assert fn == 4
    assert fv[:4].tolist() == [2.0, 3.0, 6.0, 9.0]
    # unfilled heap slots keep MaxHeap's sentinel after the reorder
    assert all(v == S for v in fv[4:].tolist())


# ---------------------------------------------------------------------------
# Independent trace checker.

def f32(x):
    return F32(x)


def pivot_of(top, qn, rr):
    """(top - |q|^2) / radius_rate / 2.0 with float operands, double divisions."""
    return f32(np.float64(f32(f32(top) - f32(qn))) / np.float64(rr) / 2.0)


def lookup_sum(codes, q, codebook, nsq, lsq):
    """Filter LUT sum of one point: entries ||q_m - cb_m[code]||^2 with
    rounded squares summed pairwise (lsq 1, 2 or 4), then four lanes over
    m mod 4 summed as (s0 + s1) + (s2 + s3)."""
    s = [f32(0)] * 4
    for m in range(nsq):
        x = q[m * lsq:(m + 1) * lsq]
        y = codebook[m, codes[m]]
        sq = [f32(f32(x[j] - y[j]) * f32(x[j] - y[j])) for j in range(lsq)]
        e = sq[0] if lsq == 1 else (f32(sq[0] + sq[1]) if lsq == 2 else
                                    f32(f32(sq[0] + sq[1]) + f32(sq[2] + sq[3])))
        s[m % 4] = f32(s[m % 4] + e)
    return f32(f32(s[0] + s[1]) + f32(s[2] + s[3]))


def check_trace(steps, fv, fn, *, cd_all, fd_all, st_all, cell_start, rec_off, rec_codes, q, codebook, qn, rr,
                scc, ftk, nsq, lsq):
    """Recomputes every decision of a walk trace.  cd_all [C], fd_all [F]
    (distinct values), st_all [C, F] (empty cells already at sqrt(FLT_MAX)),
    cell_start [C*F+1], rec_off [N], rec_codes [N, nsq] in memory order."""
    C, F = st_all.shape
    corder = np.argsort(cd_all, kind="stable")[:scc]
    forder = np.argsort(fd_all, kind="stable")
    mdo = st_all.min(axis=1)
    accepted = []

    def top():
        return f32(sorted(accepted)[ftk - 1]) if len(accepted) >= ftk else f32(S)

    si = 0
    prev_pivot = pivot_of(S, qn, rr)
    for li in range(scc):
        c = int(corder[li])
        cd = f32(cd_all[c])
        min_dist = f32(mdo[c] + cd)
        max_st = f32(f32(pivot_of(top(), qn, rr) - cd) - f32(fd_all[forder[0]]))
        for idx in range(F):
            assert si < len(steps), "trace ends early"
            t = steps[si]
            si += 1
            k = int(forder[idx])
            assert (t["l"], t["idx"], t["k"], t["coarse"]) == (li, idx, k, c), "visiting order"
            T = top()
            P = pivot_of(T, qn, rr)
            assert f32(t["top_before"]) == T, "heap top is the ftk-th smallest accepted value"
            assert f32(t["pivot_before"]) == P, "pivot = (top - |q|^2) / radius_rate / 2"
            assert P <= prev_pivot, "pivot never rises"
            prev_pivot = P
            assert f32(t["max_st_before"]) == max_st, "max_stationary_dist carried from the last scanned cell"
            fd = f32(fd_all[k])
            if f32(fd + min_dist) >= P:
                assert t["decision"] == 0, "break <=> fine_distance + min_dist >= pivot"
                break
            st = f32(st_all[c, k])
            if st >= max_st:
                assert t["decision"] == 1, "continue <=> stationary >= max_stationary_dist"
                assert f32(t["max_st_after"]) == max_st and f32(t["pivot_after"]) == P
                continue
            assert t["decision"] == 2, "scanned otherwise"
            cell_dist = f32(f32(cd + st) + fd)
            offered = updated = 0
            for i in range(int(cell_start[c * F + k]), int(cell_start[c * F + k + 1])):
                t0 = f32(2.0 * np.float64(cell_dist) + np.float64(rec_off[i]))
                if t0 >= top():
                    break
                offered += 1
                v = f32(t0 + lookup_sum(rec_codes[i], q, codebook, nsq, lsq))
                if v < top():
                    accepted.append(v)
                    updated += 1
            assert (t["offered"], t["updated"]) == (offered, updated), "points of a scanned cell"
            P2 = pivot_of(top(), qn, rr)
            assert f32(t["pivot_after"]) == P2
            max_st = min(max_st, f32(f32(P2 - cd) - fd))
            assert f32(t["max_st_after"]) == max_st, "max_stationary_dist lowered after a scanned cell"
    assert si == len(steps), "trace has cells the rules do not visit"
    want = sorted(accepted)[:ftk]
    assert fn == len(want)
    assert sorted(fv[:fn].tolist()) == [float(v) for v in want], "filter set = the ftk smallest accepted"


def check_kat_trace(name, L=None):
    k = KATS[name]
    ix, w, steps, fv, ft, fn, st = run_kat(name, L)
    ex = ix.export()
    C, F = len(k["cd"]), len(k["fd"])
    N = ex["mem2local"].shape[0]
    rec_off, rec_codes = kat_records(k)
    check_trace(steps, fv, fn, cd_all=np.asarray(k["cd"], F32), fd_all=np.asarray(k["fd"], F32),
                st_all=ex["stationary"].reshape(C, F), cell_start=ex["cell_start"], rec_off=rec_off,
                rec_codes=rec_codes, q=w["q"], codebook=np.repeat(np.arange(256, dtype=F32)[None, :, None], 2, 0),
                qn=k["qn"], rr=k["rr"], scc=k["scc"], ftk=k["ftk"], nsq=2, lsq=1)
    assert N == len(rec_off)


def kat_records(k):
    """Records of a known-answer index in memory order (cells ascending, then
    offset)."""
    F = len(k["fd"])
    rows = sorted(((c * F + kk, off, codes) for (c, kk), pts in k["cells"].items() for off, codes in pts),
                  key=lambda r: (r[0], r[1]))
    return np.array([r[1] for r in rows], F32), np.array([r[2] for r in rows], np.int64)


@pytest.mark.parametrize("name", sorted(KATS))
def test_walk_trace_checker_on_known_answers(name):
    check_kat_trace(name)


def random_walk_case(seed, L=None):
    """A random index (C = 6, F = 12, dim 8, nsq 4 of 2 floats, 900 points with
    integer-valued coordinates) and a query; walk inputs from the oracle's own
    arithmetic."""
    rng = np.random.default_rng(seed)
    C, F, d, nsq, N = 6, 12, 8, 4, 900
    coarse = rng.normal(0, 3, (C, d)).astype(F32)
    fine = rng.normal(0, 1, (F, d)).astype(F32)
    cb = rng.normal(0, 1, (nsq, 256, 2)).astype(F32)
    cell = rng.integers(0, C * F, N).astype(np.uint32)
    cell[rng.random(N) < 0.1] = 5  # a crowded cell
    recs = np.zeros((N, 4 + nsq), np.uint8)
    recs[:, :4] = np.frombuffer(rng.normal(0, 2, N).astype(F32).tobytes(), np.uint8).reshape(N, 4)
    recs[:, 4:] = rng.integers(0, 256, (N, nsq))
    raw = rng.normal(0, 1, (N, d)).astype(F32)
    ix = O.OracleIndex(coarse, fine, cb, cell, recs, raw, False, L=L)
    q = rng.normal(0, 2, d).astype(F32)
    cip = np.array([O.dot(coarse[c], q) for c in range(C)], F32)
    fip = np.array([O.dot(fine[f], q) for f in range(F)], F32)
    qn = F32(O.dot(q, q))
    return ix, dict(q=q, cip=cip, fip=fip, qn=qn, cb=cb, recs=recs, C=C, F=F, nsq=nsq)


@pytest.mark.parametrize("seed,rr,ftk,scc", [(1, 1.0, 16, 6), (2, 1.3, 8, 5), (3, 2.0, 24, 6), (4, 1.0, 3, 4),
                                             (5, 1.7, 40, 6), (6, 1.0, 64, 3)])
def test_walk_trace_checker_random(seed, rr, ftk, scc):
    ix, w = random_walk_case(seed)
    steps, fv, ft, fn, st = ix.walk_trace(w["q"], w["cip"], w["fip"], w["qn"], scc, ftk, rr)
    ex = ix.export()
    C, F, nsq = w["C"], w["F"], w["nsq"]
    m2l = ex["mem2local"]
    recs = w["recs"][m2l]
    cn = ex["coarse_norms"]
    cd = (cn - w["cip"]).astype(F32)
    fd = (-w["fip"]).astype(F32)
    assert len(set(cd.tolist())) == C and len(set(fd.tolist())) == F
    check_trace(steps, fv, fn, cd_all=cd, fd_all=fd, st_all=ex["stationary"].reshape(C, F),
                cell_start=ex["cell_start"], rec_off=recs[:, :4].copy().view(F32).reshape(-1),
                rec_codes=recs[:, 4:].astype(np.int64), q=w["q"], codebook=w["cb"], qn=w["qn"], rr=rr, scc=scc,
                ftk=ftk, nsq=nsq, lsq=2)
    # the pivot moved and pruned: some list broke before its last cell
    assert any(s["decision"] == 0 for s in steps)


# ---------------------------------------------------------------------------
# Planted mistakes: every variant must be rejected.

# variant -> the known answer built to reject it
MUTANTS = {
    "PO_MUT_PIVOT_NO_HALF": "pivot_half_and_query_norm",
    "PO_MUT_PIVOT_NO_RR": "radius_rate_and_min_dist_offset",
    "PO_MUT_PIVOT_NO_QN": "pivot_half_and_query_norm",
    "PO_MUT_BREAK_GT": "ties_at_break_and_continue",
    "PO_MUT_CONT_GT": "ties_at_break_and_continue",
    "PO_MUT_POINT_GT": "max_st_start_and_point_break",
    "PO_MUT_NO_MAXST_UPDATE": "ties_at_break_and_continue",
    "PO_MUT_MAXST_ON_CONTINUE": "ties_at_break_and_continue",
    "PO_MUT_MAXST_INIT_NO_FD0": "max_st_start_and_point_break",
    "PO_MUT_MIN_DIST_NO_MDO": "radius_rate_and_min_dist_offset",
}


def build_variant(macro, out_dir):
    src = os.path.join(os.path.dirname(O.__file__), "puck_oracle.c")
    out = os.path.join(str(out_dir), f"liboracle_{macro}.so")
    subprocess.check_call(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=c99", "-shared",
                           "-fPIC", f"-D{macro}", "-o", out, src, "-lm"])
    return O.load_variant(out)


def suite_rejects(L):
    """Names of the checks that fail for the oracle build L."""
    failed = []
    for name in sorted(KATS):
        filt, stats = kat_result(name, L)
        if filt != [float(v) for v in KATS[name]["filt"]] or stats != KATS[name]["stats"]:
            failed.append("kat:" + name)
    return failed


@pytest.mark.parametrize("macro", sorted(MUTANTS))
def test_planted_walk_mistakes_are_rejected(macro, tmp_path):
    L = build_variant(macro, tmp_path)
    failed = suite_rejects(L)
    assert "kat:" + MUTANTS[macro] in failed, f"{macro}: its known answer does not reject it ({failed})"


def test_unmutated_build_passes_the_same_harness(tmp_path):
    # the harness itself: a build with no switch passes every check
    L = build_variant("PO_NO_MUTATION", tmp_path)
    assert suite_rejects(L) == []
