"""The LDS bank model (tests/lds_bank_model.py) against measured
SQ_LDS_BANK_CONFLICT counts per wave on MI355X (rocprofv3 --pmc runs of
bench.py, recorded in profiles/iteration_k_col_k_row_512_pop64_raar.md and
profiles/iteration_block_k_iter_fused_128_pop256_raar.md)."""
import lds_bank_model as m

PLAN_512 = (8, 8, 8)


def test_access_pattern_covers_every_element_once_per_exchange():
    # each stage store and each stage load of one transform touches every
    # element of the line exactly once over the line's T threads
    for n, v, radices in ((512, 8, PLAN_512), (128, 16, (16, 8)), (384, 12, (12, 4, 4, 2))):
        T = n // v
        per_thread = [m.stage_accesses(n, v, radices, t) for t in range(T)]
        k = 0
        for s, r in enumerate(radices):
            for _ in range(2 if 0 < s < len(radices) - 1 else 1):
                touched = sorted(i for t in range(T) for i in per_thread[t][k:k + v])
                assert touched == list(range(n)), (n, s)
                k += v
        assert k == len(per_thread[0])


def test_no_conflicts_when_lanes_touch_consecutive_slots():
    assert m.access_cycles(list(range(64))) == 4
    assert m.access_cycles([0] * 64) == 4            # one address: a broadcast
    assert m.access_cycles([16 * i for i in range(64)]) == 64


def test_column_pass_512_matches_the_counter():
    assert m.column_pass(512, 8, PLAN_512, 8, pad=1) == 192      # k_col<512,0,1>: 192 measured
    assert m.column_pass(512, 8, PLAN_512, 8, skew=3) == 0       # k_col<512,0,6>: 0 measured


def test_row_pass_512_matches_the_counter():
    assert m.row_pass(512, 8, PLAN_512, 3) == 128                # k_row<512,3,0>: 128 measured
    assert m.row_pass(512, 8, PLAN_512, 3, transforms=1) == 64   # k_row<512,0,0>: 64 measured


def test_on_chip_kernel_matches_the_counter():
    # per wave and iteration: 3328 / 26 = 128 measured for the current layout
    assert m.fused_iteration(128, 136, 4, column_t_major=True) == 128
    # the first layout: 34 010 / 26 = 1308 measured; the model counts single
    # 8-byte accesses where the compiler merged some into paired ones
    old = m.fused_iteration(128, 145, 3, column_t_major=False)
    assert abs(old - 1308) / 1308 < 0.1
