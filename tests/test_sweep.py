"""Chunk-sweep harness: the JSONL rows, resume and validate semantics of the
reference's Modal launcher (modal_sym.py:36-77, 152-250), run on the local GPU."""
import json

import pytest

from map_folding_amd import sweep


def _fake_chunk(n, threads, expanded_n, start, count):
    return {"n": n, "threads": threads, "expanded_n": expanded_n, "start_case": start, "max_cases": count,
            "count": str(1000 + start), "wall_s": 0.0, "completed_hint": count, "stderr_tail": ""}


def _fake_gen(n, expanded_n, start, count):
    return ("cases", start, count), None, 0.0


def _fake_solve(n, threads, expanded_n, start, count, cases, schedule=None):
    assert cases[0] == ("cases", start, count)
    return _fake_chunk(n, threads, expanded_n, start, count)


def test_sweep_resume_validate(tmp_path, monkeypatch):
    monkeypatch.setattr(sweep, "gen_chunk", _fake_gen)
    monkeypatch.setattr(sweep, "solve_chunk", _fake_solve)
    out = tmp_path / "r.jsonl"
    logs = []
    s = sweep.sweep(n=7, expanded_n=24, total_cases=95, chunk_size=20, chunk_count=2, out=str(out), log=logs.append,
                    pipeline=False)
    assert s["chunks"] == 2 and s["partial_sum"] == str(1000 + 1020)
    rows = [json.loads(l) for l in out.read_text().splitlines()]
    assert [r["start_case"] for r in rows] == [0, 20]
    assert set(rows[0]) == {"n", "threads", "expanded_n", "start_case", "max_cases", "count", "wall_s",
                            "completed_hint", "stderr_tail"}
    assert json.loads((tmp_path / "r.jsonl.summary.json").read_text())["partial_sum"] == "2020"
    # Resume runs only the missing chunks; the last one is ragged (95 - 80 = 15 cases).
    s = sweep.sweep(n=7, expanded_n=24, total_cases=95, chunk_size=20, out=str(out), resume_existing=True,
                    log=logs.append, pipeline=False)
    assert any("skipped=2 remaining=3" in l for l in logs)
    rows = [json.loads(l) for l in out.read_text().splitlines()]
    assert [r["start_case"] for r in rows] == [0, 20, 40, 60, 80] and rows[-1]["max_cases"] == 15
    v = sweep.validate(str(out), total_cases=95, chunk_size=20, expected=str(sum(1000 + s for s in range(0, 95, 20))))
    assert v["complete"] and v["matches_expected"] and v["missing_count"] == 0
    # A duplicate row is reported and the file is no longer complete.
    with out.open("a") as f:
        f.write(json.dumps(rows[0]) + "\n")
    v = sweep.validate(str(out), total_cases=95, chunk_size=20)
    assert v["duplicate_count"] == 1 and not v["complete"] and v["unique_sum"] != v["all_rows_sum"]
    v = sweep.validate(str(out), total_cases=120, chunk_size=20)
    assert v["missing_starts"] == [100] and not v["complete"]


def test_sweep_pipeline_overlaps_generation(tmp_path, monkeypatch):
    """Chunk k+1's cases are generated while chunk k is solved, rows stay in
    chunk order, and every chunk is solved from its own cases."""
    import threading
    started = {}
    lock = threading.Lock()

    def fake_gen(n, expanded_n, start, count):
        with lock:
            started[start] = threading.Event()
        started[start].set()
        return ("cases", start, count), None, 0.0

    def fake_solve(n, threads, expanded_n, start, count, cases, schedule=None):
        assert cases[0] == ("cases", start, count)
        nxt = start + 20
        if nxt < 95:   # the next chunk's generation starts before this solve returns
            for _ in range(500):
                with lock:
                    ev = started.get(nxt)
                if ev is not None:
                    break
                threading.Event().wait(0.01)
            assert ev is not None and ev.wait(5.0), f"chunk {nxt} was not generated during chunk {start}"
        return _fake_chunk(n, threads, expanded_n, start, count)

    monkeypatch.setattr(sweep, "gen_chunk", fake_gen)
    monkeypatch.setattr(sweep, "solve_chunk", fake_solve)
    out = tmp_path / "p.jsonl"
    s = sweep.sweep(n=7, expanded_n=24, total_cases=95, chunk_size=20, out=str(out), log=lambda m: None)
    assert s["pipelined"] and s["chunks"] == 5
    rows = [json.loads(l) for l in out.read_text().splitlines()]
    assert [r["start_case"] for r in rows] == [0, 20, 40, 60, 80] and rows[-1]["max_cases"] == 15
    assert s["partial_sum"] == str(sum(1000 + k for k in range(0, 95, 20)))
    # n < 7 keeps the one-call path.
    monkeypatch.setattr(sweep, "run_chunk", _fake_chunk)
    s = sweep.sweep(n=6, expanded_n=16, total_cases=40, chunk_size=20, out=str(tmp_path / "q.jsonl"),
                    log=lambda m: None)
    assert not s["pipelined"] and s["chunks"] == 2


@pytest.mark.gpu
def test_sweep_7x7_chunk_on_gpu(gpu, golden, tmp_path):
    g = next(c for c in golden["sym_chunks"]["items"]
             if c["n"] == 7 and c["start"] == 200000 and c["max"] == 20000)
    out = tmp_path / "s.jsonl"
    s = sweep.sweep(n=7, expanded_n=g["expanded"], total_cases=433131, chunk_size=20000, start_case=200000,
                    chunk_count=1, out=str(out), log=lambda m: None)
    assert s["partial_sum"] == str(g["answer"])
    row = json.loads(out.read_text().splitlines()[0])
    assert row["count"] == str(g["answer"]) and row["start_case"] == 200000


@pytest.mark.gpu
def test_sweep_pipelined_7x7_chunks_on_gpu(gpu, golden, tmp_path):
    """Three pipelined 7x7 chunks (the middle one the golden window) add up to
    the unpipelined per-chunk answers, and the golden chunk matches."""
    g = next(c for c in golden["sym_chunks"]["items"]
             if c["n"] == 7 and c["start"] == 200000 and c["max"] == 20000)
    a = tmp_path / "a.jsonl"
    b = tmp_path / "b.jsonl"
    kw = dict(n=7, expanded_n=g["expanded"], total_cases=433131, chunk_size=20000, start_case=180000,
              chunk_count=3, log=lambda m: None)
    sa = sweep.sweep(out=str(a), pipeline=True, **kw)
    sb = sweep.sweep(out=str(b), pipeline=False, **kw)
    assert sa["pipelined"] and not sb["pipelined"]
    ra = [json.loads(l) for l in a.read_text().splitlines()]
    rb = [json.loads(l) for l in b.read_text().splitlines()]
    assert [r["count"] for r in ra] == [r["count"] for r in rb]
    assert ra[1]["start_case"] == 200000 and ra[1]["count"] == str(g["answer"])
    assert sa["partial_sum"] == sb["partial_sum"]
