"""The rocprofv3 summaries of gridcl_amd.profsum on small synthetic CSV files (CPU only)."""
import csv

from gridcl_amd import profsum

TRACE_COLS = ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X",
              "Grid_Size_Y", "Grid_Size_Z"]


def _trace_csv(tmp_path, dispatches):
    path = tmp_path / "t_kernel_trace.csv"
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(TRACE_COLS)
        t = 0
        for i, (name, grid, ns) in enumerate(dispatches, start=1):
            w.writerow([i, name, t, t + ns, grid, 1, 1])
            t += ns + 10
    return str(path)


def test_run_numbers_split_at_other_dispatches():
    rows = [{"Dispatch_Id": str(i), "k": k} for i, k in
            enumerate(["a", "a", "b", "a", "a", "c", "b", "b", "a"], start=1)]
    run = profsum.run_numbers(rows, lambda r: r["k"])
    assert [run[i] for i in range(1, 10)] == [0, 0, 0, 1, 1, 0, 1, 1, 2]


def test_trace_runs_tell_two_legs_apart(tmp_path, capsys):
    """One kernel at one grid for two point sets, a torch kernel between: --runs reports each
    leg's own average; without it the two are averaged together."""
    legs = [("k<20>(int)", 96, 100)] * 4 + [("copy", 8, 5)] + [("k<20>(int)", 96, 400)] * 4
    path = _trace_csv(tmp_path, legs)
    profsum.trace(path, "k<20>", last=0, by_run=True)
    out = capsys.readouterr().out.splitlines()[1:]
    got = {line.split()[2]: float(line.split()[5]) for line in out}
    assert got == {"1": 0.4, "0": 0.1}
    profsum.trace(path, "k<20>", last=2)
    out = capsys.readouterr().out.splitlines()[1:]
    assert len(out) == 1 and float(out[0].split()[3]) == 0.4
