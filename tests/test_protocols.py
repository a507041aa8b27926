"""The named equilibration protocols (core/src/equilibrate.cpp:43-75) as stage lists: the 21 steps of Larsen et al.
2011 (Table 1: heat at t_max, cool at t_final, compress / decompress at fractions of p_max, a long final NPT stage) and
the annealing cycles."""
import caps_amd as ca


def test_larsen21_is_table_1():
    st = ca.protocol("larsen21", ca.ProtocolParams(t_final=300, t_max=600, p_final=1, p_max=50000, time_scale=1))
    assert len(st) == 21
    ens = "".join("p" if s.ensemble == "npt" else "v" for s in st)
    assert ens == "vvp" * 7
    assert [s.ps for s in st] == [50, 50, 50, 50, 100, 50, 50, 100, 50, 50, 100, 5, 5, 10, 5, 5, 10, 5, 5, 10, 800]
    assert [s.pressure for s in st if s.ensemble == "npt"] == [1000, 30000, 50000, 25000, 5000, 500, 1]
    assert all(s.t_start == (600 if k % 3 == 0 else 300) for k, s in enumerate(st))
    half = ca.protocol("larsen21", ca.ProtocolParams(time_scale=0.5))
    assert [s.ps for s in half] == [0.5 * s.ps for s in st]


def test_annealing_cycles():
    st = ca.protocol("annealing", ca.ProtocolParams(cycles=2, t_low=300, t_high=500, ramp_ps=20, hold_ps=10, p_final=1))
    assert len(st) == 1 + 4 * 2 and all(s.ensemble == "npt" for s in st)
    assert (st[1].t_start, st[1].t_end, st[3].t_start, st[3].t_end) == (300, 500, 500, 300)
    assert [s.ps for s in st] == [10, 20, 10, 20, 10, 20, 10, 20, 10]
