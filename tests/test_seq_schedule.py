"""Forward progress of the ticket-stream schedule of the multi-workgroup
sequence sweeps (k_solve_seq, k_solve_seqp, k_solve_seqp2 in
seevr_amd/csrc/sv_pgs.hip), on the CPU.

The sweep runs the serial Gauss-Seidel order of physics.comp:500-583 (item
(t, c) = iteration t of contact c, in sweep order t * nc + c) on many
workgroups.  An item may run once every earlier item touching one of its
dynamic bodies has run (each body counts its updates: item (t, c) waits for
t * deg(a) + rank_a(c)) and the contact's previous iteration has run.
Workgroup w serves ticket stream s = w mod S, S = sv_debug_seq_streams(W);
stream s hands out, from its own counter and in order, the items of the
64-item blocks s, s + S, s + 2S, ...  A lane holds at most three items: the
current one (polled until ready, then run), the next one (promoted when the
current one is done) and a ticket (a counter value, turned into the next
item one turn later); a wave asks for tickets, with one counter atomic, for
the lanes that hold nothing, and leaves when its lanes hold nothing and
every ticket it got lies past the last item.

The model below restates that schedule and runs it with workgroups that
progress at very different rates (one nearly stalled), with fewer
workgroups than eight streams, and with 32- and 64-item waves (the paired
kernel holds 32 items per wave).  Every item must run exactly once, in an
order every body sees as the serial one, within a bounded number of turns.
A schedule with more streams than workgroups must fail the same check (a
stream without a workgroup never hands out its items)."""
import numpy as np
import pytest


def seq_streams(wgs: int) -> int:
    from seevr_amd import _lib
    return int(_lib.load().sv_debug_seq_streams(wgs))


def make_contacts(rng, n_bodies, nc, static_frac=0.1, skip_frac=0.05):
    a = rng.integers(0, n_bodies, nc)
    b = (a + 1 + rng.integers(0, n_bodies - 1, nc)) % n_bodies
    dyn = rng.random(n_bodies) >= static_frac
    skip = rng.random(nc) < skip_frac
    return a, b, dyn, skip


class Sweep:
    """Readiness and completion of items (t, c) against per-body update
    counts, the rule the kernels poll."""

    def __init__(self, a, b, dyn, skip, iters):
        self.a, self.b, self.dyn, self.skip, self.iters = a, b, dyn, skip, iters
        self.nc = a.size
        nb = dyn.size
        self.deg = np.zeros(nb, np.int64)
        self.rank_a = np.zeros(self.nc, np.int64)
        self.rank_b = np.zeros(self.nc, np.int64)
        for c in range(self.nc):  # ranks in contact order, skipped contacts are not items
            if skip[c]:
                continue
            if dyn[a[c]]:
                self.rank_a[c] = self.deg[a[c]]
                self.deg[a[c]] += 1
            if dyn[b[c]]:
                self.rank_b[c] = self.deg[b[c]]
                self.deg[b[c]] += 1
        self.count = np.zeros(nb, np.int64)
        self.citer = np.zeros(self.nc, np.int64)
        self.ran = np.zeros(iters * self.nc, np.int64)
        self.log = {k: [] for k in range(nb)}

    def ready(self, q):
        t, c = divmod(q, self.nc)
        a, b = self.a[c], self.b[c]
        return ((not self.dyn[a] or self.count[a] == t * self.deg[a] + self.rank_a[c]) and
                (not self.dyn[b] or self.count[b] == t * self.deg[b] + self.rank_b[c]) and self.citer[c] == t)

    def run(self, q):
        t, c = divmod(q, self.nc)
        for body in (self.a[c], self.b[c]):
            if self.dyn[body]:
                self.count[body] += 1
                self.log[body].append(q)
        self.citer[c] += 1
        self.ran[q] += 1


def simulate(sweep, wgs, streams, lanes_per_wave, waves_per_wg, speed, rng, max_turns):
    """Returns the turns the schedule took, or None if it did not finish."""
    total = sweep.iters * sweep.nc
    lg = streams.bit_length() - 1
    assert 1 << lg == streams
    counters = [0] * streams
    waves = []
    for w in range(wgs):
        for _ in range(waves_per_wg):
            waves.append({"wg": w, "stream": w & (streams - 1), "cur": [None] * lanes_per_wave,
                          "nxt": [None] * lanes_per_wave, "tk": [None] * lanes_per_wave,
                          "fin": [False] * lanes_per_wave, "pend": None, "done": False})
    for turn in range(max_turns):
        if all(wv["done"] for wv in waves):
            return turn
        for wv in waves:
            if wv["done"] or rng.random() >= speed[wv["wg"]]:
                continue  # this wave makes no step this turn
            cur, nxt, tk, fin = wv["cur"], wv["nxt"], wv["tk"], wv["fin"]
            for L in range(lanes_per_wave):
                if cur[L] is not None and sweep.ready(cur[L]):  # poll and apply
                    sweep.run(cur[L])
                    cur[L] = None
                if cur[L] is None and nxt[L] is not None:  # promote (a skipped contact is dropped)
                    q = nxt[L]
                    nxt[L] = None
                    if not sweep.skip[q % sweep.nc]:
                        cur[L] = q
            if wv["pend"] is not None:  # last turn's tickets
                lanes, base = wv["pend"]
                for k, L in enumerate(lanes):
                    r = base + k
                    q = ((r >> 6) << (6 + lg)) + (wv["stream"] << 6) + (r & 63)
                    if q >= total:
                        fin[L] = True
                    else:
                        tk[L] = q
                wv["pend"] = None
            for L in range(lanes_per_wave):  # a ticket becomes the next item
                if tk[L] is not None and nxt[L] is None:
                    nxt[L], tk[L] = tk[L], None
            need = [L for L in range(lanes_per_wave)
                    if cur[L] is None and nxt[L] is None and tk[L] is None and not fin[L]]
            if need:
                s = wv["stream"]
                wv["pend"] = (need, counters[s])
                counters[s] += len(need)
            if wv["pend"] is None and all(cur[L] is None and nxt[L] is None and tk[L] is None
                                          for L in range(lanes_per_wave)):
                wv["done"] = True
    return None


def check_serial(sweep):
    total = sweep.iters * sweep.nc
    items = [q for q in range(total) if not sweep.skip[q % sweep.nc]]
    assert np.all(sweep.ran[items] == 1), "every item runs exactly once"
    assert np.all(sweep.ran[[q for q in range(total) if sweep.skip[q % sweep.nc]]] == 0)
    for body, seq in sweep.log.items():
        assert seq == sorted(seq), f"body {body} saw its updates out of sweep order"


def test_stream_count_fits_the_workgroups():
    for wgs in range(1, 2000):
        s = seq_streams(wgs)
        assert s & (s - 1) == 0 and 1 <= s <= 8 and s <= wgs
    assert seq_streams(640) == 8 and seq_streams(5) == 4 and seq_streams(1) == 1


@pytest.mark.parametrize("wgs,lanes,waves", [(1, 64, 1), (2, 64, 1), (3, 32, 2), (5, 64, 1), (7, 32, 1),
                                             (8, 64, 1), (12, 32, 1), (20, 64, 1)])
def test_schedule_finishes_in_serial_order_with_skewed_workgroups(wgs, lanes, waves):
    rng = np.random.default_rng(wgs * 131 + lanes)
    a, b, dyn, skip = make_contacts(rng, n_bodies=60, nc=700)
    sweep = Sweep(a, b, dyn, skip, iters=3)
    speed = rng.uniform(0.2, 1.0, wgs)
    speed[rng.integers(0, wgs)] = 0.01  # one workgroup nearly stalled
    turns = simulate(sweep, wgs, seq_streams(wgs), lanes, waves, speed, rng, max_turns=400_000)
    assert turns is not None, "the schedule stopped making progress"
    check_serial(sweep)


def test_dense_chain_finishes():
    # every contact shares body 0: the sweep is one chain, the worst case
    rng = np.random.default_rng(2)
    nc = 300
    a = np.zeros(nc, np.int64)
    b = rng.integers(1, 40, nc)
    sweep = Sweep(a, b, np.ones(40, bool), np.zeros(nc, bool), iters=2)
    turns = simulate(sweep, 6, seq_streams(6), 64, 1, rng.uniform(0.3, 1.0, 6), rng, max_turns=200_000)
    assert turns is not None
    check_serial(sweep)


def test_more_streams_than_workgroups_stalls():
    # the check has teeth: with 8 streams served by 5 workgroups, streams 5-7
    # never hand out their blocks and the sweep cannot finish
    rng = np.random.default_rng(7)
    a, b, dyn, skip = make_contacts(rng, n_bodies=50, nc=600, skip_frac=0.0)
    sweep = Sweep(a, b, dyn, skip, iters=2)
    turns = simulate(sweep, 5, 8, 64, 1, np.ones(5), rng, max_turns=5_000)
    assert turns is None
    assert sweep.ran.sum() < 2 * 600
