"""Schedule builder and settings (host logic of the memetic loop).

The expected schedules are evaluated by hand from the reference's formula
(spring/alg_manager.py:11-16, 42-54, 105-143): weight
w(g) = clip((1 - (g - g0)/(g1 - g0))**1.5, 0, 1) with g0 = round(0.1*G),
g1 = round(0.95*G) for the default G = 100 generations (g0 = 10, g1 = 95).
"""
import numpy as np
import pytest

from s_spi_spring_amd import ER, HIO, RAAR, AlgManager, Settings


def default_manager():
    # MPR.InitIA with the default settings (spring/mpr.py:110-139)
    return AlgManager(main=[HIO(40, 0, 0.95, decay_exp=1.5), ER(40, 80, decay_exp=1.5)],
                      evaluation=[ER(40)], sw_sigma=2.5, sw_threshold=0.03, repetitions=3,
                      scalingstart=10, scalingend=95)


def test_generation_zero_sequence():
    m = default_manager()
    seq = m.get_sequence(0, area=[500, 700])
    # w(0) = clip((1 + 10/85)**1.5) = 1
    assert seq == [("HIO", [40, 0.95]), ("ER", [40]), ("SW", [2.5, 0.03]),
                   ("HIO", [40, 0.95]), ("ER", [40]), ("SWA", [2.5, 500]),
                   ("HIO", [40, 0.95]), ("ER", [40]), ("SWA", [2.5, 700.0]), ("ER", [40])]


def test_midway_generation():
    m = default_manager()
    # g = 52: w = (1 - 42/85)**1.5 = 0.359820...
    w = (1 - 42 / 85) ** 1.5
    assert abs(w - 0.35982) < 1e-5
    seq = m.get_sequence(52, area=[500, 700])
    assert seq[0] == ("HIO", [14, 0.95])            # round(0.35982*40) = 14
    assert seq[1] == ("ER", [66])                   # round(0.35982*40 + 0.64018*80) = round(65.607) = 66
    # AlgManager's shrink-wrap schedule keeps SW's default decay exponent 1 (linear):
    # w_sw = 1 - 42/85 = 0.505882..., sigma = 0.505882*2.5 + 0.494118*0.66 = 1.590824
    wsw = 1 - 42 / 85
    name, (sigma, thr) = seq[2]
    assert name == "SW"
    assert abs(sigma - (wsw * 2.5 + (1 - wsw) * 0.66)) < 1e-12 and abs(sigma - 1.590824) < 1e-6
    assert abs(thr - (wsw * 0.03 + (1 - wsw) * 0.0198)) < 1e-12
    name, (s2, area) = seq[8]
    assert name == "SWA" and abs(area - (wsw * 700 + (1 - wsw) * 500)) < 1e-9
    assert m.get_support_params(52) == [sigma, thr]


def test_after_the_decay_window():
    m = default_manager()
    # g = 100 > 95: the base is negative, the principal power's real part is ~0 -> w = 0
    seq = m.get_sequence(100, area=[500, 700])
    assert seq[0] == ("HIO", [0, 0.95]) and seq[1] == ("ER", [80])
    assert seq[2] == ("SW", [0.66, pytest.approx(0.0198)])


def test_repetitions_and_flags():
    m = AlgManager(main=[RAAR(10, beta=0.87)], evaluation=[ER(5)], repetitions=5)
    seq = m.get_sequence(0, area=[10, 10], update_support=False)
    assert seq == [("RAAR", [10, 0.87])] * 5 + [("ER", [5])]
    seq = m.get_sequence(0, area=[10, 20], update_support_first=True)
    assert seq[0][0] == "SW" and sum(1 for s in seq if s[0] == "RAAR") == 5
    assert AlgManager(main=[], repetitions=1).repetitions == 2
    ipr = m.get_sequence_ipr(0)
    assert [s[0] for s in ipr] == ["RAAR", "SW"] * 5 + ["ER"]


def test_settings_defaults_and_io(tmp_path):
    s = Settings()
    g = s.get()
    assert g["global"]["popsize"] == 128 and g["IA"]["beta"] == 0.95 and g["GA"]["crossprob"] == 0.6
    assert s.get("init", "supportsize") == 60
    assert s.get_advanced()["seed"] == 42 and s.get_advanced()["decay_exp"] == 1.5
    s.set("global", "popsize", 64).set_advanced("seed", 7)
    f = tmp_path / "s.yaml"
    s.save(str(f))
    t = Settings(str(f))
    assert t.get("global", "popsize") == 64 and t.get_advanced()["seed"] == 7
    u = Settings().setjson(s.getjson())
    assert u.get() == s.get()
