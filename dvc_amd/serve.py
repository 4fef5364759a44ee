"""Serving loop over the continuous-batching engine (dvc_engine_*).

Workload (the serving configuration): 32 concurrent synthetic streams, 24 at
QCIF and 8 at CIF, GOP drawn from {2, 4, 8} and Qindex from 1..8 per stream.
WZ-frame decode requests arrive as a Poisson process; a request is submitted
to the engine once it has arrived and its reference frames are decoded (for
GOP 4/8 the hierarchical order of decodeGOP, DVC_Decoder.cpp:471-472, makes
decoded WZ frames references of later ones).  Key frames stay resident in
the engine under per-stream keys; frames are admitted by HBM footprint and
join the running batch between bitplane steps.  Latency is measured from
arrival to completion.

    python -m dvc_amd.serve --rates 100,400,1600
"""
import argparse
import json
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import workload as W


@dataclass
class Stream:
    sid: int
    width: int
    height: int
    gop: int
    qindex: int
    video: np.ndarray
    decoded: dict = field(default_factory=dict)   # frame index -> decoded luma
    order: list = field(default_factory=list)     # (cur, past, fut) in decode order
    wz: dict = field(default_factory=dict)         # cur -> WZFrame


def gop_order(begin, end, out):
    """decodeGOP's recursion order of the WZ frames of one GOP."""
    mid = (begin + end) // 2
    if mid > begin:
        out.append((mid, begin, end))
        gop_order(begin, mid, out)
        gop_order(mid, end, out)


def make_streams(n_qcif=24, n_cif=8, frames=17, seed=4000):
    rng = np.random.default_rng(seed)
    streams = []
    for sid in range(n_qcif + n_cif):
        cif = sid >= n_qcif
        w, h = W.CIF if cif else W.QCIF
        gop = int(rng.choice([2, 4, 8]))
        q = int(rng.integers(1, 9))
        vid = W.synthetic_video(w, h, frames, 6 if cif else 3, seed + sid)
        st = Stream(sid, w, h, gop, q, vid)
        for k in range(0, frames - 1, gop):
            end = min(k + gop, frames - 1)
            st.decoded[k] = vid[k]
            st.decoded[end] = vid[end]
            gop_order(k, end, st.order)
        for cur, _, _ in st.order:
            st.wz[cur] = W.encode_wz_frame(vid[cur], q)
        streams.append(st)
    return streams


def key_id(sid, frame):
    return ((sid + 1) << 32) | frame


def run(ctx, streams, rate, seed=1, max_inflight=64, lanes=0, duration_cap=120.0, verify=None):
    """Feeds Poisson arrivals at `rate` requests/s and steps the engine.
    Returns a dict of throughput and latency figures."""
    rng = np.random.default_rng(seed)
    # request list: each stream's WZ frames in decode order, interleaved by arrival
    reqs = []
    cursor = {s.sid: 0 for s in streams}
    total = sum(len(s.order) for s in streams)
    t = 0.0
    live = [s for s in streams]
    while len(reqs) < total:
        t += rng.exponential(1.0 / rate)
        cand = [s for s in live if cursor[s.sid] < len(s.order)]
        s = cand[int(rng.integers(0, len(cand)))]
        cur, past, fut = s.order[cursor[s.sid]]
        cursor[s.sid] += 1
        reqs.append(dict(t=t, s=s, cur=cur, past=past, fut=fut))
    eng = _lib.Engine(ctx, max_inflight=max_inflight, lanes=lanes)
    t0 = time.perf_counter()
    nxt = 0
    waiting = []        # arrived, references not decoded yet
    tickets = {}
    lat = []
    steps = 0
    done = 0
    outputs = {}
    while done < total:
        now = time.perf_counter() - t0
        while nxt < total and reqs[nxt]["t"] <= now:
            waiting.append(reqs[nxt])
            nxt += 1
        still = []
        for r in waiting:
            s = r["s"]
            if r["past"] in s.decoded and r["fut"] in s.decoded:
                fr = dict(past=s.decoded[r["past"]], fut=s.decoded[r["fut"]], orig=s.video[r["cur"]],
                          wz=s.wz[r["cur"]], idx=(r["past"], r["cur"], r["fut"]))
                tk = eng.submit(fr, key_id(s.sid, r["past"]), key_id(s.sid, r["fut"]))
                tickets[tk] = r
            else:
                still.append(r)
        waiting = still
        if not tickets:
            if nxt < total:
                time.sleep(max(0.0, min(reqs[nxt]["t"] - now, 0.01)))
            continue
        eng.step()
        steps += 1
        fin_t = time.perf_counter() - t0
        for tk in list(tickets):
            if eng.done(tk):
                r = tickets.pop(tk)
                rec, st = eng.result(tk)
                r["s"].decoded[r["cur"]] = rec
                outputs[(r["s"].sid, r["cur"])] = (rec, st)
                lat.append(fin_t - r["t"])
                done += 1
        if fin_t > duration_cap:
            break
    wall = time.perf_counter() - t0
    eng.close()
    res = dict(rate=rate, requests=total, completed=done, wall_s=wall, steps=steps,
               frames_per_s=done / wall if wall > 0 else 0.0,
               latency_ms=dict(p50=float(np.percentile(lat, 50) * 1e3) if lat else None,
                               p99=float(np.percentile(lat, 99) * 1e3) if lat else None,
                               mean=float(np.mean(lat) * 1e3) if lat else None))
    if verify is not None:
        res["verified"] = verify(outputs, streams)
    return res, outputs


def decode_gop(eng, video, wz, begin, end, sid=0):
    """Decodes the WZ frames of decodeGOP(begin, end) (DVC_Decoder.cpp:26-474)
    through engine `eng`.  A frame is submitted as soon as both of its
    references are decoded, so the frames of one level of the hierarchy run
    together; the key frames are video[begin] and video[end].  Returns
    {cur: (recons, stats)} and the reference's decode order."""
    order = []
    gop_order(begin, end, order)
    decoded = {begin: video[begin], end: video[end]}
    todo, tickets, out = list(order), {}, {}
    while todo or tickets:
        for item in list(todo):
            cur, past, fut = item
            if past in decoded and fut in decoded:
                fr = dict(past=decoded[past], fut=decoded[fut], orig=video[cur], wz=wz[cur], idx=(past, cur, fut))
                tickets[eng.submit(fr, key_id(sid, past), key_id(sid, fut))] = cur
                todo.remove(item)
        eng.step()
        for tk, cur in list(tickets.items()):
            if eng.done(tk):
                rec, st = eng.result(tk)
                decoded[cur] = rec
                out[cur] = (rec, st)
                del tickets[tk]
    return out, order


def references(stream, cur):
    """(past, fut) of WZ frame `cur` of a stream."""
    return [(p, f) for c, p, f in stream.order if c == cur][0]


def verify_solo(ctx, sample=8, seed=0):
    """Checks a sample of served frames (all of them when sample <= 0)
    against decodes of the same frame alone (an engine with one frame in
    flight), with the references the serving loop used."""
    def check(outputs, streams):
        rng = np.random.default_rng(seed)
        keys = list(outputs)
        if sample <= 0:
            pick = keys
        else:
            pick = [keys[i] for i in rng.choice(len(keys), min(sample, len(keys)), replace=False)]
        bad = 0
        eng = _lib.Engine(ctx, max_inflight=1, lanes=1)
        try:
            for sid, cur in pick:
                s = streams[sid]
                past, fut = references(s, cur)
                fr = dict(past=s.decoded[past], fut=s.decoded[fut], orig=s.video[cur], wz=s.wz[cur],
                          idx=(past, cur, fut))
                (rec, st), = eng.decode_all([fr])
                srec, sst = outputs[(sid, cur)]
                bad += int(not (np.array_equal(rec, srec) and st["levels"] == sst["levels"]
                                and st["psnr"] == sst["psnr"]))
        finally:
            eng.close()
        return dict(checked=len(pick), mismatches=bad)
    return check


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="100,400,1600")
    ap.add_argument("--frames", type=int, default=17)
    ap.add_argument("--verify", type=int, default=8)
    args = ap.parse_args()
    ctx = _lib.Context()
    for rate in [float(r) for r in args.rates.split(",")]:
        streams = make_streams(frames=args.frames)
        res, _ = run(ctx, streams, rate, verify=verify_solo(ctx, args.verify) if args.verify else None)
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
