"""The kernel-trace summary behind the lane-timing cross-check: busy time is
the union of a kernel's dispatch intervals, so overlapping dispatches on
several streams count once.  CPU only."""
from acoustic_isotropic_operators_amd.profiling import summarize


def test_busy_time_is_the_union_of_dispatch_intervals(tmp_path):
    p = tmp_path / "run_kernel_trace.csv"
    p.write_text('"Kind","Kernel_Name","Start_Timestamp","End_Timestamp"\n'
                 '"KERNEL_DISPATCH","void k<32, 16>(A)",100,200\n'
                 '"KERNEL_DISPATCH","void k<32, 16>(A)",150,260\n'
                 '"KERNEL_DISPATCH","other",0,1000\n'
                 '"KERNEL_DISPATCH","void k<32, 16>(A)",300,340\n')
    s = summarize(str(p), ["k<"])
    assert list(s) == ["void k<32, 16>(A)"]
    k = s["void k<32, 16>(A)"]
    assert k["dispatches"] == 3
    assert k["sum_ns"] == 100 + 110 + 40
    assert k["busy_ns"] == 160 + 40            # [100, 260) and [300, 340)
    assert k["span_ns"] == 240
