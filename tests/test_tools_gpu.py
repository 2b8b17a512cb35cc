"""The scorer micro-benchmark (python -m hybridsearchsystem_amd.scorebench) runs end to end on a small
store: every dtype it names, a batch that takes the staged scorer (BF16 384 dims at 128 queries) and one
that takes the LDS-query-tile scorer (64 queries), with and without survivors."""
import pytest

pytest.importorskip("torch")


@pytest.mark.gpu
def test_scorebench_runs(capsys):
    from hybridsearchsystem_amd import scorebench
    print("scorebench small run", flush=True)
    rc = scorebench.main(["--docs", "70000", "--dim", "384", "--dtypes", "f32,bf16,fp8,f16x2", "--queries", "64,128",
                          "--floors", "inf,0.05", "--reps", "2"])
    assert rc == 0
    out = capsys.readouterr().out
    lines = [l for l in out.splitlines() if " queries " in l and " floor " in l]
    assert len(lines) == 4 * 2 * 2
    # a finite floor lets survivors through; +inf lets none
    assert all(l.endswith("survivors/query 0") for l in lines if "floor   inf" in l)
    assert any(not l.endswith("survivors/query 0") for l in lines if "floor  0.05" in l)
