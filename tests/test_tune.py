"""The launch-shape autotuner: its store, modes and text file on the CPU; on the GPU, runs
with the cost model, with tuning (every valid form of each new shape class measured), with
the frozen store, and with every measured form's output compared with the first form's,
give the same bits as the oracle."""
import numpy as np
import pytest


def _read(path):
    with open(path) as f:
        return f.read()


def test_tune_store_round_trip(tmp_path):
    from lakcluster_c_amd import lkc
    before = lkc.tune_get()["mode"]
    try:
        lkc.tune_clear()
        f = tmp_path / "tune.txt"
        f.write_text("lkc-tune 1 20 18 12 3 4 2 0.5 0.75\n"
                     "lkc-tune 2 14 16 5 0 0 1 0.1 0.1\n"
                     "not a tuning line\n")
        assert lkc.tune_load(str(f)) == 2
        assert lkc.tune_get()["classes"] == 2
        g = tmp_path / "again.txt"
        assert lkc.tune_save(str(g)) == 2
        lines = sorted(g.read_text().split("\n"))
        assert lines[1:] == ["lkc-tune 1 20 18 12 3 4 2 0.5 0.75", "lkc-tune 2 14 16 5 0 0 1 0.1 0.1"]
        for m in (0, 1, 2, 3):
            lkc.tune_set(m)
            assert lkc.tune_get()["mode"] == m
        with pytest.raises(lkc.LkcError):
            lkc.tune_set(4)
        with pytest.raises(lkc.LkcError):
            lkc.tune_load(str(tmp_path / "missing.txt"))
        lkc.tune_clear()
        assert lkc.tune_get()["classes"] == 0
    finally:
        lkc.tune_set(before)


@pytest.mark.gpu
@pytest.mark.parametrize("kw", [
    dict(nA=700, ncols=5000, nZ=600, T=3, gamma=0.05, B=32, seed=61),
    dict(nA=400, ncols=9000, nZ=0, T=2, gamma=0.08, B=24, seed=62),
])
def test_tuned_forms_match_oracle(kw, oracle, tmp_path):
    from lakcluster_c_amd import gen, lkc
    ds = gen.make_dataset(str(tmp_path), "tune", **kw)
    text = _read(ds.in_path)
    print(f"tuner parity: {kw}", flush=True)
    ref = oracle.run_text(text)
    before = lkc.tune_get()["mode"]
    runs = {}
    try:
        for label, mode in (("model", 0), ("measure", 1), ("frozen", 2), ("verify", 3)):
            if mode in (1, 3):
                lkc.tune_clear()
            lkc.tune_set(mode)
            with lkc.Bcc(text) as run:
                run.run()
                runs[label] = (run.trace(), run.xdrop())
            if mode == 1:
                st = lkc.tune_get()
                print(f"  measured {st['measured']} classes, saved {st['saved_ms']:.3f} ms", flush=True)
                assert st["measured"] > 0
                f = str(tmp_path / "store.txt")
                n = lkc.tune_save(f)
                lkc.tune_clear()
                assert lkc.tune_load(f) == n
            if mode == 3:
                # every measured form wrote the same bits as the first form of its launch
                compared, mismatched = lkc.tune_verify()
                print(f"  verify: {compared} launches compared, {mismatched} words differed", flush=True)
                assert compared > 0
                assert mismatched == 0
    finally:
        lkc.tune_set(before)
    for label, (tr, (xa, xb)) in runs.items():
        assert tr.tobytes() == ref.trace.tobytes(), label
        np.testing.assert_array_equal(xa, ref.xdrop_a, err_msg=label)
        np.testing.assert_array_equal(xb, ref.xdrop_b, err_msg=label)
